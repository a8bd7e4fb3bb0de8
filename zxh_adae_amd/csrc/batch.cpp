// batch.cpp -- batched parameterized circuits (see batch.h).
#include "batch.h"

#include <algorithm>
#include <cstring>

namespace zxa
{

namespace
{
size_t align256(size_t b)
{
    return (b + 255) & ~size_t(255);
}

std::vector<unsigned> targets_of(const std::vector<U3Gate> &gs)
{
    std::vector<unsigned> t;
    for (const U3Gate &g : gs)
        t.push_back(unsigned(__builtin_ctzll(g.delta)));
    return t;
}
} // namespace

AnsatzBatch::AnsatzBatch(int prec) : prec_(prec)
{
}

AnsatzBatch::~AnsatzBatch()
{
    if (ctx_.stream)
        (void)hipStreamSynchronize(ctx_.stream);
    for (int b = 0; b < 2; b++)
    {
        if (exec_[b])
            (void)hipGraphExecDestroy(exec_[b]);
        if (graph_[b])
            (void)hipGraphDestroy(graph_[b]);
    }
    fwd_.clear();
    bwd_.clear();
    init_.reset();
    obs_fwd_.reset();
    obs_bwd_.reset();
    for (void *p : {arena_, static_cast<void *>(d_grads_), static_cast<void *>(d_energy_),
                    static_cast<void *>(d_states_)})
        if (p)
            (void)hipFree(p);
}

int AnsatzBatch::add(unsigned n, const AnsatzOp *ops, size_t nops, size_t nparams, const int32_t *qi,
                     const int32_t *qj, const double *coef, size_t nterms)
{
    if (final_)
    {
        set_error("batch: add after finalize");
        return ESTATE;
    }
    Set s;
    s.a.reset(new Ansatz(n, prec_, nparams));
    if (n < 65)
        s.a->set_geometry(geometry_[n]);
    ZXA_TRY(s.a->init(ops, nops, false));
    ZXA_TRY(s.a->set_observable(qi, qj, coef, nterms));
    s.poff = total_params_;
    total_params_ += nparams;
    sets_.push_back(std::move(s));
    return OK;
}

int AnsatzBatch::set_geometry(unsigned n, unsigned seg)
{
    if (n >= 65 || (seg != 0 && (seg < 3 || seg > 6)))
    {
        set_error("batch: geometry must be 0 or 3..6, width at most 64");
        return EINVAL_;
    }
    if (final_)
    {
        set_error("batch: set_geometry after finalize");
        return ESTATE;
    }
    geometry_[n] = seg;
    return OK;
}

int AnsatzBatch::finalize()
{
    if (final_)
        return OK;
    if (sets_.empty())
    {
        set_error("batch: no sets");
        return EINVAL_;
    }
    ZXA_TRY(ctx_.init());
    // arena: every psi, then every lam
    size_t off = 0;
    std::vector<size_t> offs;
    for (Set &s : sets_)
    {
        offs.push_back(off);
        off += align256((size_t(1) << s.a->nbits()) * elem_bytes(prec_));
    }
    psi_bytes_ = off;
    ZXA_HIP(hipMalloc(&arena_, 2 * psi_bytes_));
    ZXA_HIP(hipMalloc(reinterpret_cast<void **>(&d_grads_), align256(std::max<size_t>(1, total_params_) * 8)));
    ZXA_HIP(hipMalloc(reinterpret_cast<void **>(&d_energy_), align256(2 * sets_.size() * 8)));
    ZXA_HIP(hipMalloc(reinterpret_cast<void **>(&d_states_), sets_.size() * sizeof(uint64_t)));
    std::vector<uint64_t> ptrs;
    for (size_t i = 0; i < sets_.size(); i++)
    {
        Set &s = sets_[i];
        s.psi = static_cast<char *>(arena_) + offs[i];
        s.lam = static_cast<char *>(arena_) + psi_bytes_ + offs[i];
        s.grads = d_grads_ + s.poff;
        ptrs.push_back(reinterpret_cast<uint64_t>(s.psi));
    }
    ZXA_HIP(hipMemcpy(d_states_, ptrs.data(), ptrs.size() * sizeof(uint64_t), hipMemcpyHostToDevice));

    size_t maxp = 0;
    for (Set &s : sets_)
        maxp = std::max(maxp, s.a->tile_passes());
    std::vector<U3Gate> gs;
    std::vector<DiagGate> w;
    for (size_t p = 0; p < maxp; p++)
    {
        std::vector<BatchPassSpec> specs;
        std::vector<std::pair<size_t, size_t>> map;
        for (size_t i = 0; i < sets_.size(); i++)
        {
            Set &s = sets_[i];
            if (s.a->tile_passes() <= p)
                continue;
            const std::vector<double> zero(std::max<size_t>(1, s.a->nparams()), 0.0);
            s.a->pass_gates(p, zero.data(), gs);
            s.a->pass_window(p, zero.data(), false, w);
            BatchPassSpec sp;
            sp.psi = s.psi;
            sp.n = s.a->nbits();
            sp.geometry = int(s.a->geometry());
            sp.targets = targets_of(gs);
            sp.window = w;
            specs.push_back(sp);
            map.push_back({i, p});
        }
        fwd_.emplace_back(new BatchLaunch());
        ZXA_TRY(batch_prepare(ctx_, prec_, false, specs, *fwd_.back()));
        fwd_map_.push_back(map);
    }
    // backward step r runs pass P_s - 1 - r of every set with P_s > r; its
    // epilogue is the previous pass's window with negated angles
    for (size_t r = 0; r < maxp; r++)
    {
        std::vector<BatchPassSpec> specs;
        std::vector<std::pair<size_t, size_t>> map;
        for (size_t i = 0; i < sets_.size(); i++)
        {
            Set &s = sets_[i];
            const size_t P = s.a->tile_passes();
            if (P <= r)
                continue;
            const size_t j = P - 1 - r;
            const std::vector<double> zero(std::max<size_t>(1, s.a->nparams()), 0.0);
            s.a->pass_gates(j, zero.data(), gs);
            BatchPassSpec sp;
            sp.psi = s.psi;
            sp.lam = s.lam;
            sp.n = s.a->nbits();
            sp.geometry = int(s.a->geometry());
            sp.targets = targets_of(gs);
            if (j > 0)
                s.a->pass_window(j - 1, zero.data(), true, sp.window);
            sp.grads = s.grads;
            sp.nparams = s.a->nparams();
            specs.push_back(sp);
            map.push_back({i, j});
        }
        bwd_.emplace_back(new BatchLaunch());
        ZXA_TRY(batch_prepare(ctx_, prec_, true, specs, *bwd_.back()));
        bwd_map_.push_back(map);
    }
    // the last pass's window of every set, undone on both states first
    {
        std::vector<BatchPassSpec> specs;
        for (size_t i = 0; i < sets_.size(); i++)
        {
            Set &s = sets_[i];
            const size_t P = s.a->tile_passes();
            if (P == 0)
                continue;
            const std::vector<double> zero(std::max<size_t>(1, s.a->nparams()), 0.0);
            BatchPassSpec sp;
            s.a->pass_window(P - 1, zero.data(), true, sp.window);
            if (sp.window.empty())
                continue;
            sp.psi = s.psi;
            sp.lam = s.lam;
            sp.n = s.a->nbits();
            sp.geometry = int(s.a->geometry());
            sp.grads = s.grads;
            sp.nparams = s.a->nparams();
            specs.push_back(sp);
            init_map_.push_back(i);
        }
        if (!specs.empty())
        {
            init_.reset(new BatchLaunch());
            ZXA_TRY(batch_prepare(ctx_, prec_, true, specs, *init_));
        }
    }
    std::vector<ObsSpec> of, ob;
    for (Set &s : sets_)
    {
        ObsSpec o;
        o.src = s.psi;
        o.n = s.a->nbits();
        o.qi = s.a->obs_qi();
        o.qj = s.a->obs_qj();
        o.coef = s.a->obs_coef();
        of.push_back(o);
        o.dst = s.lam;
        ob.push_back(o);
    }
    obs_fwd_.reset(new ObsBatch());
    obs_bwd_.reset(new ObsBatch());
    ZXA_TRY(obs_batch_prepare(ctx_, prec_, of, d_energy_, *obs_fwd_));
    ZXA_TRY(obs_batch_prepare(ctx_, prec_, ob, d_energy_, *obs_bwd_));
    ZXA_HIP(hipStreamSynchronize(ctx_.stream));
    ZXA_TRY(capture(false));
    ZXA_TRY(capture(true));
    final_ = true;
    return OK;
}

// The launch sequence of one run (kernels only: fills instead of memset
// nodes), enqueued directly or recorded into a graph.
int AnsatzBatch::enqueue(bool backward)
{
    int rc = zero_fill(ctx_, arena_, psi_bytes_);
    if (!rc)
        rc = set_ones(ctx_, d_states_, int(sets_.size()), prec_);
    for (size_t p = 0; p < fwd_.size() && !rc; p++)
        rc = batch_launch(ctx_, *fwd_[p]);
    if (!rc && backward)
    {
        rc = zero_fill(ctx_, d_grads_, align256(std::max<size_t>(1, total_params_) * 8));
        if (!rc)
            rc = obs_batch_launch(ctx_, *obs_bwd_);
        if (!rc && init_)
            rc = batch_launch(ctx_, *init_);
        for (size_t r = 0; r < bwd_.size() && !rc; r++)
            rc = batch_launch(ctx_, *bwd_[r]);
    }
    else if (!rc)
    {
        rc = obs_batch_launch(ctx_, *obs_fwd_);
    }
    return rc;
}

int AnsatzBatch::capture(bool backward)
{
    const int b = backward ? 1 : 0;
    // one eager run first: the runtime sets up every kernel (code objects,
    // private segments) outside the capture
    ZXA_TRY(enqueue(backward));
    ZXA_HIP(hipStreamSynchronize(ctx_.stream));
    ZXA_HIP(hipStreamBeginCapture(ctx_.stream, hipStreamCaptureModeRelaxed));
    const int rc = enqueue(backward);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx_.stream, &g);
    if (rc || e != hipSuccess)
    {
        if (g)
            (void)hipGraphDestroy(g);
        if (rc)
            return rc;
        set_error(std::string("batch: graph capture failed: ") + hipGetErrorString(e));
        return EHIP;
    }
    graph_[b] = g;
    ZXA_HIP(hipGraphInstantiate(&exec_[b], g, nullptr, nullptr, 0));
    return OK;
}

int AnsatzBatch::fill_tables(const double *params, bool backward)
{
    std::vector<U3Gate> gs;
    for (size_t p = 0; p < fwd_.size(); p++)
    {
        for (size_t k = 0; k < fwd_map_[p].size(); k++)
        {
            const Set &s = sets_[fwd_map_[p][k].first];
            s.a->pass_gates(fwd_map_[p][k].second, params + s.poff, gs);
            batch_set_gates(*fwd_[p], k, gs.data(), gs.size());
        }
        ZXA_TRY(batch_upload(ctx_, *fwd_[p]));
    }
    if (!backward)
        return OK;
    std::vector<AdjTarget> ts;
    for (size_t r = 0; r < bwd_.size(); r++)
    {
        for (size_t k = 0; k < bwd_map_[r].size(); k++)
        {
            const Set &s = sets_[bwd_map_[r][k].first];
            s.a->pass_adjoint(bwd_map_[r][k].second, params + s.poff, ts);
            ZXA_TRY(batch_set_adjoint(*bwd_[r], k, ts.data(), ts.size()));
        }
        ZXA_TRY(batch_upload(ctx_, *bwd_[r]));
    }
    if (init_)
        ZXA_TRY(batch_upload(ctx_, *init_));
    return OK;
}

int AnsatzBatch::run(const double *params, double *energies, double *grads)
{
    if (!final_)
    {
        set_error("batch: run before finalize");
        return ESTATE;
    }
    const bool backward = grads != nullptr;
    ZXA_TRY(fill_tables(params, backward));
    ZXA_HIP(hipGraphLaunch(exec_[backward ? 1 : 0], ctx_.stream));
    if (energies)
        ZXA_HIP(hipMemcpyAsync(energies, d_energy_, 2 * sets_.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx_.stream));
    if (backward && total_params_)
        ZXA_HIP(hipMemcpyAsync(grads, d_grads_, total_params_ * sizeof(double), hipMemcpyDeviceToHost, ctx_.stream));
    return hip_check(hipStreamSynchronize(ctx_.stream), "batch run");
}

} // namespace zxa
