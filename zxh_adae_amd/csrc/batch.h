// batch.h -- many parameterized circuits of different widths and depths
// (expectation values and adjoint gradients) evaluated together.  No
// reference counterpart.
//
// Every set keeps its two states (psi, lam) and its gradient array in one
// device arena.  The sets' schedules (ansatz.h) are merged by pass index:
// forward pass p of every set that has one runs in ONE persistent launch of
// the batched tile kernel (work item = one tile of one set), then one batched
// observable launch gives every set's energy and lam = H psi, then one batched
// adjoint launch per backward step.  The whole sequence is captured once in a
// hipGraph and replayed; per run only the matrix tables (a pinned mirror) are
// uploaded before the replay.  Each set's tile passes use that set's own tile
// geometry (the ansatz schedule's).
#pragma once

#include "ansatz.h"

#include <memory>
#include <vector>

namespace zxa
{

class AnsatzBatch
{
  public:
    explicit AnsatzBatch(int prec);
    ~AnsatzBatch();
    int add(unsigned n, const AnsatzOp *ops, size_t nops, size_t nparams, const int32_t *qi, const int32_t *qj,
            const double *coef, size_t nterms);
    int finalize();
    // Tile geometry (3..6; 0: the fewest-passes rule) for the sets of n
    // qubits added after this call.
    int set_geometry(unsigned n, unsigned seg);
    // params: every set's parameters, concatenated in set order.  energies
    // [2 * sets]: <H>, <psi|psi> per set.  grads (nullable: forward only):
    // concatenated like params.
    int run(const double *params, double *energies, double *grads);
    size_t sets() const
    {
        return sets_.size();
    }
    size_t total_params() const
    {
        return total_params_;
    }
    unsigned set_nbits(size_t i) const
    {
        return sets_[i].a->nbits();
    }
    size_t set_passes(size_t i) const
    {
        return sets_[i].a->tile_passes();
    }
    size_t launches() const
    {
        return fwd_.size() + bwd_.size() + (init_ ? 1 : 0) + 2;
    }
    Ctx &ctx()
    {
        return ctx_;
    }

  private:
    unsigned geometry_[65] = {}; // per width: forced tile geometry (0: the rule)
    struct Set
    {
        std::unique_ptr<Ansatz> a;
        size_t poff = 0; // offset into the concatenated parameters
        void *psi = nullptr;
        void *lam = nullptr;
        double *grads = nullptr;
    };
    int enqueue(bool backward);
    int capture(bool backward);
    int fill_tables(const double *params, bool backward);

    int prec_;
    bool final_ = false;
    std::vector<Set> sets_;
    size_t total_params_ = 0;
    void *arena_ = nullptr;
    size_t psi_bytes_ = 0; // all psi states, contiguous at the arena's start
    double *d_grads_ = nullptr;
    double *d_energy_ = nullptr;
    uint64_t *d_states_ = nullptr; // psi pointers, for the reset
    std::vector<std::unique_ptr<BatchLaunch>> fwd_, bwd_;
    std::unique_ptr<BatchLaunch> init_;
    std::unique_ptr<ObsBatch> obs_fwd_, obs_bwd_;
    // per launch, per item: (set, pass) it carries
    std::vector<std::vector<std::pair<size_t, size_t>>> fwd_map_, bwd_map_;
    std::vector<size_t> init_map_;
    hipGraph_t graph_[2] = {nullptr, nullptr};
    hipGraphExec_t exec_[2] = {nullptr, nullptr};
    Ctx ctx_;
};

} // namespace zxa
