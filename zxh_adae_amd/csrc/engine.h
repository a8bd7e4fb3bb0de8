// engine.h -- host scheduler of the MI355X engine: the ZXH simulator object.
//
// Same semantics as the reference's ZXH class (src/zxh/include/zxhsim/zxh.h,
// src/zxh/src/core/zxh.cpp): gates are recorded, then execute() replays them
// over an affine address map psi_virt[A x_r ^ b] = psi_phys[x_r] with lazy
// state growth, folds consecutive H/U3 gates into one 2x2 matrix per qubit
// (dropping products within 1e-5 of the identity), and batches Z/Rz/P/CP
// into diagonal windows, flushing both exactly where the reference does.
// The flushed batches run as HIP passes (kernels_*.hip) on one GPU.
#pragma once

#include "dev.h"

#include <complex>
#include <cstdint>
#include <vector>

namespace zxa
{

enum GateCode
{
    G_NONE = 0,
    G_BARRIER = 1,
    G_Z = 2,
    G_RZ = 3,
    G_P = 4,
    G_CRZ = 5,
    G_CP = 6,
    G_X = 7,
    G_CX = 8,
    G_H = 9,
    G_U3 = 10,
    G_RX = 11
};

enum Policy
{
    POLICY_REFERENCE = 0, // every flushed batch is its own pass, as in the reference
    POLICY_FUSED = 1      // flushed batches merge into tile passes when exact
};

struct GateRec
{
    int type;
    size_t cq, q;
    double theta, lambda, phi;
};

class Sim
{
  public:
    Sim(size_t n, bool disable_x, bool eager, int prec);
    ~Sim();
    int init(); // device context

    int add_gate(int code, size_t cq, size_t q, double theta, double lambda, double phi);
    void set_seed(uint64_t s)
    {
        use_seed_ = true;
        seed_ = s;
    }
    void clear_seed()
    {
        use_seed_ = false;
        seed_ = 0;
    }
    void clear_gates()
    {
        gates_.clear();
        gphase_ = {1.0, 0.0};
    }
    int execute();
    int measure(size_t shots);
    size_t measured_count() const
    {
        return result_count_;
    }
    int get_results(uint64_t *packed, size_t shots) const;
    int get_real_results(uint64_t *real, size_t shots) const;
    int get_state(double *out, size_t out_len) const;
    size_t required_M() const;
    size_t num_qubits() const
    {
        return N_;
    }
    void set_seg_bits(size_t b)
    {
        seg_min_i_ = b ? b : 18;
    }
    int get_phys_state(void *out, size_t out_bytes, size_t *m_out) const;
    int set_phys_state(const void *in, size_t in_bytes);
    int device_ptr(void **p, size_t *m) const;
    Ctx &ctx()
    {
        return ctx_;
    }
    int precision() const
    {
        return prec_;
    }
    void set_policy(int p)
    {
        policy_ = p;
    }
    int policy() const
    {
        return policy_;
    }

  private:
    struct Acc
    {
        bool active = false;
        std::complex<double> u00{1.0, 0.0}, u01{0.0, 0.0}, u10{0.0, 0.0}, u11{1.0, 0.0};
        // a product hoisted from a later run keeps the descriptor of its own
        // place in the circuit (else it is taken from the map at the flush)
        bool fixed = false;
        uint64_t sel = 0, delta = 0;
    };

    bool solve(size_t q, uint64_t *x) const;
    uint64_t local_selector(size_t q) const;
    unsigned active_bits() const
    {
        return static_cast<unsigned>(used_bits_);
    }
    int expand();
    uint64_t solve_expand(size_t q, int *rc);
    int flush_diag(std::vector<DiagGate> &window);
    int flush_u3(std::vector<Acc> &pend, size_t &count);
    int emit_u3(std::vector<U3Gate> &descs);
    int flush_pass();
    bool fusable() const;
    int resize_reset(size_t M);
    // The reset of execute() is left pending under the fused policy: the
    // first tile pass forms |0...0> itself (fused_pass_apply's fresh input);
    // any other first use of the state runs the reset (settle).
    int settle();

    size_t N_;
    bool disable_x_, eager_;
    int prec_;
    std::vector<GateRec> gates_;
    bool use_seed_ = false;
    uint64_t seed_ = 0;
    std::complex<double> gphase_{1.0, 0.0};
    // affine map: rows_[q] bit c = A[q][c]; b_[q]
    std::vector<uint64_t> rows_;
    std::vector<uint8_t> b_;
    size_t m_ = 0;
    // physical state
    void *dstate_ = nullptr;
    bool fresh_ = false;      // the state is |0...0> of fresh_elems_ amplitudes, not yet written
    size_t fresh_elems_ = 0;
    size_t alloc_elems_ = 0;
    size_t M_ = 0;
    size_t used_bits_ = 0;
    bool executed_ = false;
    size_t seg_min_i_ = 18;
    // measure cache
    // pinned host buffers of the last measure: thresholds, real addresses and
    // (unless the map is the identity, when they equal the real addresses)
    // packed virtual words
    PinBuf thr_pin_, real_pin_, packed_pin_;
    bool packed_is_real_ = true;
    size_t result_count_ = 0;
    // pending fused pass: single-bit U3 gates (each after its fan, if any),
    // then a diagonal window
    int policy_ = POLICY_FUSED;
    std::vector<U3Gate> pass_u3_;
    std::vector<GateFan> pass_fan_;
    std::vector<DiagGate> pass_diag_;
    uint64_t pass_support_ = 0;
    int pass_high_ = 0;
    bool pass_has_fan_ = false;
    bool take_fan(unsigned t, GateFan &fan);
    // a cluster of U3 gates with wide pair offsets or per-pair selectors
    // (after CX transport), held back until the next flush so that a
    // diagonal window flushed right after it can ride its last tile pass
    // (u3_apply with the window as a table epilogue)
    std::vector<U3Gate> wide_u3_;
    int flush_wide(std::vector<DiagGate> *window = nullptr);
    // expansions in a row while nothing else is pending, applied at once
    // before the next physical operation
    std::vector<U3Gate> exp_group_;
    unsigned exp_base_ = 0;
    int flush_group();
    Ctx ctx_;
};

// Runtime flag of init()/finalize() (runtime_general.cpp:22-49).
bool runtime_active();
void runtime_set_active(bool on);

} // namespace zxa
