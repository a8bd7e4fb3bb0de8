// ansatz.cpp -- schedule, forward and adjoint backward of parameterized
// circuits (see ansatz.h).
#include "ansatz.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace zxa
{

namespace
{
bool single_qubit(int kind)
{
    return kind == A_RX || kind == A_RY || kind == A_RZ || kind == A_H;
}
bool rotation(int kind)
{
    return kind == A_RX || kind == A_RY || kind == A_RZ;
}
} // namespace

Ansatz::Ansatz(unsigned n, int prec, size_t nparams) : n_(n), prec_(prec), nparams_(nparams)
{
}

Ansatz::~Ansatz()
{
    if (ctx_.stream)
        (void)hipStreamSynchronize(ctx_.stream);
    if (psi_)
        (void)hipFree(psi_);
    if (lam_)
        (void)hipFree(lam_);
    if (dgrads_)
        (void)hipFree(dgrads_);
}

int Ansatz::init(const AnsatzOp *ops, size_t nops, bool device)
{
    if (n_ < tile_min_bits(prec_) || n_ > 36)
    {
        set_error("ansatz: qubit count must be at least one tile (12 for complex128, 13 for complex64) and at "
                  "most 36");
        return EINVAL_;
    }
    for (size_t i = 0; i < nops; i++)
    {
        const AnsatzOp &o = ops[i];
        const bool two = o.kind == A_CZ || o.kind == A_CP;
        if (o.kind < A_RX || o.kind > A_CP)
        {
            set_error("ansatz: unknown operation kind");
            return EINVAL_;
        }
        if (o.q0 < 0 || unsigned(o.q0) >= n_ || (two && (o.q1 < 0 || unsigned(o.q1) >= n_ || o.q1 == o.q0)))
        {
            set_error("ansatz: qubit index outside the circuit");
            return EINVAL_;
        }
        if (o.param >= 0 && (size_t(o.param) >= nparams_ || !rotation(o.kind)))
        {
            set_error("ansatz: parameter index out of range or on a gate without a parameter");
            return EINVAL_;
        }
    }
    ops_.assign(ops, ops + nops);
    ZXA_TRY(build_schedule());
    if (!device)
        return OK;
    ZXA_TRY(ctx_.init());
    return alloc();
}


int Ansatz::alloc()
{
    const size_t bytes = (size_t(1) << n_) * elem_bytes(prec_);
    hipError_t e = hipMalloc(&psi_, bytes);
    if (e == hipSuccess)
        e = hipMalloc(&lam_, bytes);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&dgrads_), std::max<size_t>(1, nparams_) * sizeof(double));
    if (e != hipSuccess)
    {
        set_error(std::string("ansatz: device allocation failed: ") + hipGetErrorString(e));
        return ENOMEM_;
    }
    return OK;
}

int Ansatz::set_observable(const int32_t *qi, const int32_t *qj, const double *coef, size_t nterms)
{
    for (size_t k = 0; k < nterms; k++)
        if (qi[k] < 0 || unsigned(qi[k]) >= n_ || qj[k] >= int32_t(n_))
        {
            set_error("ansatz: observable qubit outside the circuit");
            return EINVAL_;
        }
    qi_.assign(qi, qi + nterms);
    qj_.assign(qj, qj + nterms);
    coef_.assign(coef, coef + nterms);
    return OK;
}

// The schedule depends only on the structure: the tile geometry with the
// fewest passes, ties going to the longest contiguous runs, unless a
// geometry was set (the batch autotuner's timed choice).
int Ansatz::build_schedule()
{
    std::vector<Pass> best, p;
    unsigned bseg = 0;
    if (force_seg_ >= 3 && force_seg_ <= 6)
    {
        schedule_for(force_seg_, best);
        passes_.swap(best);
        seg_ = force_seg_;
        tile_passes_ = passes_.size();
        return OK;
    }
    for (unsigned seg = 6; seg >= 3; seg--)
    {
        schedule_for(seg, p);
        if (bseg == 0 || p.size() < best.size())
        {
            best.swap(p);
            bseg = seg;
        }
    }
    passes_.swap(best);
    seg_ = bseg;
    tile_passes_ = passes_.size();
    return OK;
}

// Passes of the circuit under tile geometry seg.
void Ansatz::schedule_for(unsigned seg, std::vector<Pass> &passes_) const
{
    passes_.clear();
    const unsigned low = tile_low_bits(prec_, int(seg));
    const unsigned slots = tile_upper_slots(prec_, int(seg));
    std::vector<Target> pending(n_);
    std::vector<int> nrot(n_, 0);
    std::vector<size_t> window;
    uint64_t support = 0;

    auto close = [&]() {
        std::vector<Target> lows, highs;
        for (unsigned q = 0; q < n_; q++)
            if (!pending[q].ops.empty())
            {
                pending[q].q = q;
                (q < low ? lows : highs).push_back(pending[q]);
            }
        if (lows.empty() && highs.empty() && window.empty())
            return;
        size_t hi = 0;
        bool first = true;
        while (first || hi < highs.size())
        {
            Pass p;
            if (first)
                p.targets = lows;
            const size_t take = std::min<size_t>(slots, highs.size() - hi);
            for (size_t k = 0; k < take; k++)
                p.targets.push_back(highs[hi + k]);
            hi += take;
            first = false;
            passes_.push_back(p);
        }
        passes_.back().window = window;
        for (unsigned q = 0; q < n_; q++)
        {
            pending[q].ops.clear();
            nrot[q] = 0;
        }
        window.clear();
        support = 0;
    };

    for (size_t i = 0; i < ops_.size(); i++)
    {
        const AnsatzOp &o = ops_[i];
        if (single_qubit(o.kind))
        {
            const unsigned q = unsigned(o.q0);
            if ((support >> q) & 1u)
                close();
            if (o.param >= 0 && nrot[q] == 2)
                close();
            pending[q].ops.push_back(i);
            if (o.param >= 0)
                nrot[q]++;
        }
        else
        {
            window.push_back(i);
            support |= (uint64_t(1) << o.q0) | (uint64_t(1) << o.q1);
        }
    }
    close();
}

Ansatz::Mat Ansatz::op_matrix(size_t i, const double *params) const
{
    const AnsatzOp &o = ops_[i];
    const double th = o.param >= 0 ? params[o.param] : o.value;
    const double c = std::cos(th / 2), s = std::sin(th / 2);
    switch (o.kind)
    {
    case A_RX:
        return {c, {0, -s}, {0, -s}, c};
    case A_RY:
        return {c, -s, s, c};
    case A_RZ:
        return {std::polar(1.0, -th / 2), 0.0, 0.0, std::polar(1.0, th / 2)};
    default: // A_H
    {
        const double h = 1.0 / std::sqrt(2.0);
        return {h, h, h, -h};
    }
    }
}

namespace
{
using cplx = std::complex<double>;
struct M2
{
    cplx a, b, c, d;
};
M2 mul(const M2 &x, const M2 &y)
{
    return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
M2 dag(const M2 &x)
{
    return {std::conj(x.a), std::conj(x.c), std::conj(x.b), std::conj(x.d)};
}
void put(const M2 &m, double *u)
{
    const cplx e[4] = {m.a, m.b, m.c, m.d};
    for (int k = 0; k < 4; k++)
    {
        u[2 * k] = e[k].real();
        u[2 * k + 1] = e[k].imag();
    }
}
M2 generator(int kind)
{
    if (kind == A_RX)
        return {0.0, 1.0, 1.0, 0.0};
    if (kind == A_RY)
        return {0.0, cplx(0, -1), cplx(0, 1), 0.0};
    return {1.0, 0.0, 0.0, -1.0};
}
} // namespace

void Ansatz::pass_gates(size_t pi, const double *params, std::vector<U3Gate> &gs) const
{
    gs.clear();
    for (const Target &t : passes_[pi].targets)
    {
        M2 u{1.0, 0.0, 0.0, 1.0};
        for (size_t i : t.ops)
        {
            const Mat m = op_matrix(i, params);
            u = mul({m.a, m.b, m.c, m.d}, u);
        }
        U3Gate g;
        g.packed_sel = uint64_t(1) << t.q;
        g.delta = uint64_t(1) << t.q;
        put(u, g.u);
        gs.push_back(g);
    }
}

void Ansatz::pass_window(size_t pi, const double *params, bool negate, std::vector<DiagGate> &w) const
{
    w.clear();
    for (size_t i : passes_[pi].window)
    {
        const AnsatzOp &o = ops_[i];
        const double th = o.kind == A_CZ ? M_PI : (o.param >= 0 ? params[o.param] : o.value);
        w.push_back({uint64_t(1) << o.q0, uint64_t(1) << o.q1, negate ? -th : th});
    }
}

void Ansatz::pass_adjoint(size_t pi, const double *params, std::vector<AdjTarget> &ts) const
{
    ts.clear();
    for (const Target &t : passes_[pi].targets)
    {
        std::vector<M2> ms;
        for (size_t i : t.ops)
        {
            const Mat m = op_matrix(i, params);
            ms.push_back({m.a, m.b, m.c, m.d});
        }
        M2 u{1.0, 0.0, 0.0, 1.0};
        for (const M2 &m : ms)
            u = mul(m, u);
        AdjTarget a;
        std::memset(&a, 0, sizeof(a));
        a.target = t.q;
        put(dag(u), a.m);
        a.param[0] = a.param[1] = -1;
        int slot = 0;
        for (size_t k = 0; k < t.ops.size(); k++)
        {
            const AnsatzOp &o = ops_[t.ops[k]];
            if (o.param < 0)
                continue;
            M2 v{1.0, 0.0, 0.0, 1.0}; // gates composed after op k
            for (size_t l = k + 1; l < ms.size(); l++)
                v = mul(ms[l], v);
            put(mul(mul(v, generator(o.kind)), dag(v)), a.a[slot]);
            a.param[slot] = o.param;
            slot++;
        }
        ts.push_back(a);
    }
}

int Ansatz::run_forward(const double *params)
{
    ZXA_TRY(state_reset(ctx_, psi_, prec_, size_t(1) << n_));
    std::vector<U3Gate> gs;
    std::vector<DiagGate> w;
    for (size_t p = 0; p < passes_.size(); p++)
    {
        pass_gates(p, params, gs);
        pass_window(p, params, false, w);
        ZXA_TRY(fused_pass_apply(ctx_, psi_, prec_, n_, gs.data(), gs.size(), w.data(), w.size(), int(seg_)));
    }
    return OK;
}

int Ansatz::forward(const double *params, double *energy)
{
    state_valid_ = false;
    ZXA_TRY(run_forward(params));
    double e2[2] = {0.0, 0.0};
    ZXA_TRY(observable_apply(ctx_, psi_, nullptr, prec_, n_, qi_.data(), qj_.data(), coef_.data(), coef_.size(),
                             e2));
    if (energy)
    {
        energy[0] = e2[0];
        energy[1] = e2[1];
    }
    last_params_.assign(params, params + nparams_);
    state_valid_ = true;
    return OK;
}

int Ansatz::backward(const double *params, double *grads, double *energy)
{
    const bool same = state_valid_ && std::equal(last_params_.begin(), last_params_.end(), params);
    state_valid_ = false;
    if (!same)
        ZXA_TRY(run_forward(params));
    double e2[2] = {0.0, 0.0};
    ZXA_TRY(observable_apply(ctx_, psi_, lam_, prec_, n_, qi_.data(), qj_.data(), coef_.data(), coef_.size(), e2));
    if (energy)
    {
        energy[0] = e2[0];
        energy[1] = e2[1];
    }
    ZXA_HIP(hipMemsetAsync(dgrads_, 0, std::max<size_t>(1, nparams_) * sizeof(double), ctx_.stream));
    std::vector<DiagGate> w;
    if (!passes_.empty() && !passes_.back().window.empty())
    {
        pass_window(passes_.size() - 1, params, true, w);
        ZXA_TRY(diag_apply(ctx_, psi_, prec_, n_, w.data(), w.size()));
        ZXA_TRY(diag_apply(ctx_, lam_, prec_, n_, w.data(), w.size()));
    }
    std::vector<AdjTarget> ts;
    for (size_t jj = passes_.size(); jj-- > 0;)
    {
        pass_adjoint(jj, params, ts);
        w.clear();
        if (jj > 0)
            pass_window(jj - 1, params, true, w);
        ZXA_TRY(adjoint_pass_apply(ctx_, psi_, lam_, prec_, n_, ts.data(), ts.size(), w.data(), w.size(), dgrads_,
                                   nparams_, int(seg_)));
    }
    if (nparams_ > 0)
        ZXA_HIP(hipMemcpyAsync(grads, dgrads_, nparams_ * sizeof(double), hipMemcpyDeviceToHost, ctx_.stream));
    ZXA_HIP(hipStreamSynchronize(ctx_.stream));
    return OK;
}

} // namespace zxa
