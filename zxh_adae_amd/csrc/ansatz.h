// ansatz.h -- parameterized circuits with expectation values and adjoint
// gradients (no reference counterpart: the reference has no gradient path).
//
// A circuit is a list of operations on n qubits (plain addressing: qubit q is
// address bit q): parameterized or fixed RX / RY / RZ, H, and the diagonal
// two-qubit CZ and CP.  The schedule is built once from the structure:
// single-qubit operations compose into one 2x2 matrix per qubit, diagonal
// operations gather into a window, and a pass (U3 gates, then the window)
// closes when a single-qubit operation meets a qubit of the pending window,
// or when a qubit already carries two parameterized rotations.  A pass whose
// targets exceed one tile's slots splits into several tile passes; the
// window rides on the last.  The tile geometry is the circuit's own: the one
// with the fewest tile passes, ties going to the longest contiguous runs.
//
// forward(params): |0> -> psi through the tile passes, then <psi|H|psi> for
// the diagonal Z / ZZ observable (FP64 compensated).
// backward(params): the adjoint method without stored intermediate states:
// lam = H psi; then, pass by pass in reverse, one dual-state tile pass
// accumulates the transition values Im <lam| A |psi> of every rotation of
// the pass (A = V G V^dagger, G the rotation's generator and V the
// single-qubit gates composed after it) and uncomputes the pass on both
// states, with the previous pass's window (negated angles) as its epilogue.
// d<H>/d(theta) = Im <lam_k| G |psi_k> for a gate exp(-i theta G / 2).
#pragma once

#include "dev.h"

#include <complex>
#include <cstdint>
#include <vector>

namespace zxa
{

enum AnsatzKind
{
    A_RX = 0,
    A_RY = 1,
    A_RZ = 2,
    A_H = 3,
    A_CZ = 4,
    A_CP = 5
};

struct AnsatzOp
{
    int32_t kind;
    int32_t q0, q1; // q1: second qubit of CZ / CP
    int32_t param;  // parameter index, or < 0 for the fixed angle value
    double value;
};

class Ansatz
{
  public:
    Ansatz(unsigned n, int prec, size_t nparams);
    ~Ansatz();
    // device = false: schedule only (no stream, no device states), for the
    // batched engine, which keeps every set's states in one arena.
    int init(const AnsatzOp *ops, size_t nops, bool device = true);
    // Before init: build the schedule for tile geometry seg (3..6) instead of
    // the one with the fewest passes (0 restores that rule).
    void set_geometry(unsigned seg)
    {
        force_seg_ = seg;
    }
    int set_observable(const int32_t *qi, const int32_t *qj, const double *coef, size_t nterms);
    // energy[0] = <psi|H|psi>, energy[1] = <psi|psi>
    int forward(const double *params, double *energy);
    int backward(const double *params, double *grads, double *energy);
    void *state() const
    {
        return psi_;
    }
    unsigned nbits() const
    {
        return n_;
    }
    size_t tile_passes() const
    {
        return tile_passes_;
    }
    Ctx &ctx()
    {
        return ctx_;
    }
    int precision() const
    {
        return prec_;
    }
    size_t nparams() const
    {
        return nparams_;
    }
    // Tile geometry of the schedule (3..6 contiguous unit-address bits).
    unsigned geometry() const
    {
        return seg_;
    }
    // Pass p of the forward schedule: its U3 gates (fused matrices at params)
    // and its diagonal window (angles negated when negate).
    void pass_gates(size_t p, const double *params, std::vector<U3Gate> &gates) const;
    void pass_window(size_t p, const double *params, bool negate, std::vector<DiagGate> &window) const;
    // The adjoint targets of pass p (uncompute matrices, transition operators).
    void pass_adjoint(size_t p, const double *params, std::vector<AdjTarget> &targets) const;
    const std::vector<int32_t> &obs_qi() const
    {
        return qi_;
    }
    const std::vector<int32_t> &obs_qj() const
    {
        return qj_;
    }
    const std::vector<double> &obs_coef() const
    {
        return coef_;
    }

  private:
    using cplx = std::complex<double>;
    struct Mat
    {
        cplx a, b, c, d; // [[a, b], [c, d]]
    };
    struct Target
    {
        unsigned q;
        std::vector<size_t> ops; // composed in order (first applied first)
    };
    struct Pass
    {
        std::vector<Target> targets;
        std::vector<size_t> window; // CZ / CP operations
    };
    Mat op_matrix(size_t i, const double *params) const;
    int build_schedule();
    void schedule_for(unsigned seg, std::vector<Pass> &passes) const;
    int run_forward(const double *params);
    int alloc();

    unsigned n_;
    int prec_;
    size_t nparams_;
    std::vector<AnsatzOp> ops_;
    std::vector<Pass> passes_; // tile passes, in forward order
    size_t tile_passes_ = 0;
    unsigned seg_ = 0;       // tile geometry the schedule was built for
    unsigned force_seg_ = 0; // geometry asked for by set_geometry (0: the rule)
    std::vector<int32_t> qi_, qj_;
    std::vector<double> coef_;
    void *psi_ = nullptr;
    void *lam_ = nullptr;
    double *dgrads_ = nullptr;
    std::vector<double> last_params_;
    bool state_valid_ = false;
    Ctx ctx_;
};

} // namespace zxa
