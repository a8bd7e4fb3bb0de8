// engine.cpp -- the ZXH simulator object over HIP passes.
//
// Host algorithm restated from the reference's src/zxh/src/core/zxh.cpp
// (execute 312-553, calc_mem 648-681, solve_expand 701-719, expand_all
// 721-729, apply_x/apply_cx 731-780, apply_u3 782-825, get_state 579-636),
// src/zxh/src/core/bitmat.cpp (GF(2) solve) and
// src/zxh/src/measure/measure_common.cpp (measure 73-92).  The affine map is
// held as N row masks of at most 63 physical columns.
#include "engine.h"
#include "poly.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>

namespace zxa
{

namespace
{
constexpr double kPi = 3.14159265358979323846;
// kU3BatchEps is the float literal 1e-5f widened to double (zxh.cpp:29).
const double kU3Eps = static_cast<double>(1e-5f);
std::atomic<bool> g_active{false};

std::complex<double> phase(double a)
{
    return std::exp(std::complex<double>(0.0, a));
}

uint64_t low_mask(size_t nbits)
{
    if (nbits == 0)
        return 0;
    if (nbits >= 64)
        return ~uint64_t(0);
    return (uint64_t(1) << nbits) - 1;
}

uint64_t pack_sel(uint64_t alpha, bool beta)
{
    return (alpha & ((uint64_t(1) << 63) - 1)) | (beta ? (uint64_t(1) << 63) : 0);
}

// U3 matrix of the reference (zxh.cpp:101-113).
void u3_matrix(double theta, double lambda, double phi, std::complex<double> u[4])
{
    const double c = std::cos(theta / 2.0);
    const double s = std::sin(theta / 2.0);
    const std::complex<double> el = phase(lambda), ep = phase(phi), epl = phase(phi + lambda);
    u[0] = std::complex<double>(c, 0.0);
    u[1] = -el * s;
    u[2] = ep * s;
    u[3] = epl * c;
}

bool close_val(std::complex<double> a, std::complex<double> b)
{
    return std::abs(a - b) <= kU3Eps;
}

// GF(2) solve A x = e_q by Gauss-Jordan on N x m rows (bitmat.cpp:46-122).
// A always has full column rank, so the solution is unique when it exists.
bool gf2_solve(const std::vector<uint64_t> &rows, size_t m, size_t q, uint64_t *x)
{
    const size_t N = rows.size();
    if (m == 0)
    {
        *x = 0;
        return false;
    }
    std::vector<uint64_t> r(rows);
    std::vector<uint8_t> rhs(N, 0);
    rhs[q] = 1;
    int pivot_row[64];
    for (size_t c = 0; c < m; c++)
        pivot_row[c] = -1;
    size_t row = 0;
    for (size_t c = 0; c < m && row < N; c++)
    {
        size_t piv = row;
        while (piv < N && !((r[piv] >> c) & 1u))
            piv++;
        if (piv == N)
            continue;
        std::swap(r[piv], r[row]);
        std::swap(rhs[piv], rhs[row]);
        for (size_t k = 0; k < N; k++)
            if (k != row && ((r[k] >> c) & 1u))
            {
                r[k] ^= r[row];
                rhs[k] ^= rhs[row];
            }
        pivot_row[c] = int(row);
        row++;
    }
    for (size_t k = 0; k < N; k++)
        if (r[k] == 0 && rhs[k])
            return false;
    uint64_t sol = 0;
    for (size_t c = 0; c < m; c++)
        if (pivot_row[c] >= 0 && rhs[size_t(pivot_row[c])])
            sol |= uint64_t(1) << c;
    *x = sol;
    return true;
}

// The reference's selector-class rule for diagonal buckets
// (kernel_cuda.cu:222-232): a CP whose selector mixes lane and slice bits is
// not batched; it flushes the window and runs alone.
int sel_class(uint64_t packed)
{
    const uint64_t a = packed & ((uint64_t(1) << 63) - 1);
    if ((a >> 8) == 0)
        return 0;
    if ((a & 0xffu) == 0)
        return 1;
    return 2;
}
} // namespace

bool runtime_active()
{
    return g_active.load();
}
void runtime_set_active(bool on)
{
    g_active.store(on);
}

// ZXHSIM_SEG_MIN_I as parse_env_size_t reads it (sv_common.cpp:49-66):
// decimal strtoull of the whole string; unset, empty, trailing text or 0
// give the default of 18.  The reference reads it once per process; each
// Sim reads it at construction, which gives the same value to any process
// that sets it before its first simulator.
static size_t env_seg_min_i()
{
    const char *env = std::getenv("ZXHSIM_SEG_MIN_I");
    if (env == nullptr || *env == '\0')
        return 18;
    char *end = nullptr;
    const unsigned long long v = std::strtoull(env, &end, 10);
    if (end == env || (end != nullptr && *end != '\0') || v == 0)
        return 18;
    return static_cast<size_t>(v);
}

Sim::Sim(size_t n, bool disable_x, bool eager, int prec)
    : N_(n), disable_x_(disable_x), eager_(eager), prec_(prec), rows_(n, 0), b_(n, 0), seg_min_i_(env_seg_min_i())
{
}

Sim::~Sim()
{
    if (dstate_)
    {
        (void)hipStreamSynchronize(ctx_.stream);
        (void)hipFree(dstate_);
    }
}

int Sim::init()
{
    return ctx_.init();
}

// Recording rules of ZXH::Barrier..Rx (zxh.cpp:234-310).
int Sim::add_gate(int code, size_t cq, size_t q, double theta, double lambda, double phi)
{
    auto bad = [](const char *m) {
        set_error(m);
        return int(EINVAL_);
    };
    switch (code)
    {
    case G_BARRIER:
        gates_.push_back({G_BARRIER, 0, 0, 0.0, 0.0, 0.0});
        return OK;
    case G_RZ:
    case G_P:
    case G_Z:
    case G_X:
    case G_H:
        if (q >= N_)
            return bad("target qubit out of range");
        gates_.push_back({code, 0, q, (code == G_RZ || code == G_P) ? theta : 0.0, 0.0, 0.0});
        return OK;
    case G_U3:
        if (q >= N_)
            return bad("U3 target qubit out of range");
        gates_.push_back({G_U3, 0, q, theta, lambda, phi});
        return OK;
    case G_RX:
        if (q >= N_)
            return bad("Rx target qubit out of range");
        gates_.push_back({G_U3, 0, q, theta, kPi / 2.0, -kPi / 2.0});
        return OK;
    case G_CRZ:
        if (cq >= N_ || q >= N_)
            return bad("CRz qubit out of range");
        gates_.push_back({G_P, 0, cq, -0.5 * theta, 0.0, 0.0});
        gates_.push_back({G_CP, cq, q, theta, 0.0, 0.0});
        return OK;
    case G_CP:
        if (cq >= N_ || q >= N_)
            return bad("CP qubit out of range");
        gates_.push_back({G_CP, cq, q, theta, 0.0, 0.0});
        return OK;
    case G_CX:
        if (cq >= N_ || q >= N_)
            return bad("CX qubit out of range");
        gates_.push_back({G_CX, cq, q, 0.0, 0.0, 0.0});
        return OK;
    default:
        return bad("unknown gate code");
    }
}

bool Sim::solve(size_t q, uint64_t *x) const
{
    return gf2_solve(rows_, m_, q, x);
}

// make_selector + make_local_selector with block_start = 0 (zxh.cpp:55-60, 683-691)
uint64_t Sim::local_selector(size_t q) const
{
    return pack_sel(rows_[q] & low_mask(active_bits()), b_[q] != 0);
}

// calc_mem (zxh.cpp:648-681)
size_t Sim::required_M() const
{
    std::vector<uint64_t> rows(N_, 0);
    size_t m = 0;
    for (const GateRec &g : gates_)
    {
        bool needs = false;
        if (g.type == G_X || g.type == G_CX)
        {
            if (!disable_x_)
            {
                if (g.type == G_CX)
                    rows[g.q] ^= rows[g.cq];
            }
            else
                needs = true;
        }
        else if (g.type == G_H || g.type == G_U3)
            needs = true;
        if (needs)
        {
            uint64_t d;
            if (!gf2_solve(rows, m, g.q, &d))
            {
                if (m >= 63)
                    return 64; // beyond any state the engine can hold
                rows[g.q] |= uint64_t(1) << m;
                m++;
            }
        }
    }
    return m;
}

// sv_t::resize + reset (sv_common.cpp:286-328): the buffer only grows; the
// first 2^M amplitudes are zeroed and psi[0] = 1.
int Sim::resize_reset(size_t M)
{
    if (M > 40)
    {
        set_error("state needs more than 2^40 amplitudes");
        return ENOMEM_;
    }
    const size_t cap = size_t(1) << M;
    if (alloc_elems_ < cap)
    {
        if (dstate_)
        {
            ZXA_HIP(hipStreamSynchronize(ctx_.stream));
            ZXA_HIP(hipFree(dstate_));
            dstate_ = nullptr;
            alloc_elems_ = 0;
        }
        hipError_t e = hipMalloc(&dstate_, cap * elem_bytes(prec_));
        if (e != hipSuccess)
        {
            dstate_ = nullptr;
            set_error(std::string("state allocation failed: ") + hipGetErrorString(e));
            return ENOMEM_;
        }
        alloc_elems_ = cap;
    }
    M_ = M;
    used_bits_ = 0;
    if (policy_ == POLICY_FUSED)
    {
        fresh_ = true;
        fresh_elems_ = cap;
        return OK;
    }
    fresh_ = false;
    return state_reset(ctx_, dstate_, prec_, cap);
}

int Sim::settle()
{
    if (!fresh_)
        return OK;
    fresh_ = false;
    return state_reset(ctx_, dstate_, prec_, fresh_elems_);
}

int Sim::expand()
{
    if (used_bits_ >= M_)
    {
        set_error("expand exceeds max_bits");
        return ESTATE;
    }
    used_bits_++;
    return OK;
}

// solve_expand (zxh.cpp:701-719)
uint64_t Sim::solve_expand(size_t k, int *rc)
{
    *rc = OK;
    uint64_t delta = 0;
    if (!solve(k, &delta))
    {
        *rc = expand();
        if (*rc)
            return 0;
        rows_[k] |= uint64_t(1) << m_;
        m_++;
        delta = uint64_t(1) << (m_ - 1);
    }
    return delta;
}

int Sim::flush_diag(std::vector<DiagGate> &window)
{
    if (window.empty())
        return OK;
    ZXA_TRY(flush_group());
    if (!wide_u3_.empty())
        return flush_wide(&window);
    int rc = OK;
    bool ride = fusable() && window_is_polynomial(active_bits(), window.data(), window.size()) &&
                window_is_polynomial(active_bits(), pass_diag_.data(), pass_diag_.size());
    if (fusable() && !ride && !pass_has_fan_ && unsigned(pass_high_) <= tile_upper_slots(prec_))
    {
        // selectors beyond the polynomial form: the pass's window as a table
        std::vector<DiagGate> all = pass_diag_;
        all.insert(all.end(), window.begin(), window.end());
        ride = window_tabulable(active_bits(), all.data(), all.size());
    }
    if (ride)
    {
        // the window follows every U3 of the pending pass: it becomes (part
        // of) the pass's diagonal epilogue
        const uint64_t amask = (uint64_t(1) << 63) - 1;
        for (const DiagGate &g : window)
        {
            pass_diag_.push_back(g);
            pass_support_ |= g.s0 & amask;
            if (g.s1 != kNoSel)
                pass_support_ |= g.s1 & amask;
        }
    }
    else
    {
        rc = flush_pass();
        if (!rc)
            rc = settle();
        if (!rc)
            rc = diag_apply(ctx_, dstate_, prec_, active_bits(), window.data(), window.size());
    }
    window.clear();
    return rc;
}

bool Sim::fusable() const
{
    return policy_ == POLICY_FUSED && used_bits_ >= tile_min_bits(prec_);
}

// Runs the pending fused pass: its U3 gates, then its diagonal window.
int Sim::flush_group()
{
    if (exp_group_.empty())
        return OK;
    ZXA_TRY(settle());
    const int rc = u3_expand_group_apply(ctx_, dstate_, prec_, exp_base_, exp_group_.data(), exp_group_.size());
    exp_group_.clear();
    return rc;
}

// The held-back wide cluster, with the window (which follows it in the
// circuit, cleared here) as its last pass's epilogue when it is tabulable,
// else in a sweep of its own.
int Sim::flush_wide(std::vector<DiagGate> *window)
{
    std::vector<U3Gate> descs;
    descs.swap(wide_u3_);
    int rc = descs.empty() && !(window && !window->empty()) ? OK : settle();
    if (rc)
        return rc;
    if (window && !window->empty())
    {
        rc = u3_apply(ctx_, dstate_, prec_, active_bits(), descs.data(), descs.size(), window->data(), window->size());
        window->clear();
    }
    else if (!descs.empty())
        rc = u3_apply(ctx_, dstate_, prec_, active_bits(), descs.data(), descs.size());
    return rc;
}

int Sim::flush_pass()
{
    ZXA_TRY(flush_wide());
    if (pass_u3_.empty() && pass_diag_.empty())
        return OK;
    ZXA_TRY(flush_group());
    // the first pass after the reset forms |0...0> itself when it covers the
    // whole reset state (it resets the state where its kernel cannot)
    const bool fresh = fresh_ && (size_t(1) << active_bits()) == fresh_elems_;
    if (!fresh)
        ZXA_TRY(settle());
    fresh_ = false;
    int rc = fused_pass_apply(ctx_, dstate_, prec_, active_bits(), pass_u3_.data(), pass_u3_.size(),
                              pass_diag_.data(), pass_diag_.size(), -1,
                              pass_has_fan_ ? pass_fan_.data() : nullptr, fresh);
    pass_u3_.clear();
    pass_fan_.clear();
    pass_diag_.clear();
    pass_support_ = 0;
    pass_high_ = 0;
    pass_has_fan_ = false;
    return rc;
}

// Moves the pending window's terms that involve address bit t into a fan
// for the next gate on t (applied just before it), when they are all of
// the form theta [x_t] L with L = 1, a constant selector or [x_k ^ beta]
// (theta x_t (beta + (1 - 2 beta) x_k)), or one two-bit selector
// theta [x_t ^ x_k ^ beta] (its part without x_t stays in the window).
// False (nothing moved) otherwise.
bool Sim::take_fan(unsigned t, GateFan &fan)
{
    // fans ride passes whose window is a phase polynomial
    if (!window_is_polynomial(active_bits(), pass_diag_.data(), pass_diag_.size()))
        return false;
    const uint64_t amask = (uint64_t(1) << 63) - 1;
    const uint64_t bit = uint64_t(1) << t;
    const unsigned n = active_bits();
    fan = GateFan();
    fan.theta.assign(n, 0.0);
    std::vector<DiagGate> keep;
    for (const DiagGate &d : pass_diag_)
    {
        const uint64_t a0 = d.s0 & amask;
        const uint64_t a1 = d.s1 == kNoSel ? 0 : (d.s1 & amask);
        if (!((a0 | a1) & bit))
        {
            keep.push_back(d);
            continue;
        }
        if (d.s1 == kNoSel && a0 != bit)
        {
            // one selector on t and k: theta [x_t ^ x_k ^ beta] = the fan
            // theta (1 - 2 beta) x_t (1 - 2 x_k) plus theta [x_k ^ beta]
            const uint64_t ak = a0 & ~bit;
            if ((ak & (ak - 1)) != 0 || (ak >> n) != 0)
                return false;
            const double sg = (d.s0 >> 63) ? -1.0 : 1.0;
            fan.c += d.theta * sg;
            fan.theta[size_t(__builtin_ctzll(ak))] += -2.0 * d.theta * sg;
            keep.push_back({ak | (d.s0 & ~amask), kNoSel, d.theta});
            continue;
        }
        // the selector on t, and the other factor
        const bool on0 = (a0 & bit) != 0;
        const uint64_t st = on0 ? d.s0 : d.s1;
        if ((st & amask) != bit || (st >> 63) != 0)
            return false;
        if (d.s1 == kNoSel)
        {
            fan.c += d.theta;
            continue;
        }
        const uint64_t so = on0 ? d.s1 : d.s0;
        const uint64_t ao = so & amask;
        const double beta = (so >> 63) ? 1.0 : 0.0;
        if (ao == bit || ao == 0)
        {
            // [x_t][x_t] = x_t; a constant selector is its beta
            fan.c += ao == bit ? (beta == 0.0 ? d.theta : 0.0) : d.theta * beta;
            continue;
        }
        if ((ao & (ao - 1)) != 0 || (ao >> n) != 0)
            return false;
        fan.c += d.theta * beta;
        fan.theta[size_t(__builtin_ctzll(ao))] += d.theta * (1.0 - 2.0 * beta);
    }
    fan.on = true;
    pass_diag_.swap(keep);
    pass_support_ = 0;
    for (const DiagGate &d : pass_diag_)
    {
        pass_support_ |= d.s0 & amask;
        if (d.s1 != kNoSel)
            pass_support_ |= d.s1 & amask;
    }
    return true;
}

// Adds single-bit U3 batches to the pending pass when exact: a gate joins
// only if it commutes with the pass's diagonal window (its target is outside
// the window's support), and the tile has room (targets below the tile's low
// bits always fit, at most tile_upper_slots others).  A second gate on a pass target is composed with
// the first (applied after it).
int Sim::emit_u3(std::vector<U3Gate> &descs)
{
    bool simple = fusable();
    for (const U3Gate &g : descs)
    {
        const uint64_t alpha = g.packed_sel & ((uint64_t(1) << 63) - 1);
        simple = simple && g.delta != 0 && (g.delta & (g.delta - 1)) == 0 && alpha == g.delta;
    }
    if (!simple)
    {
        ZXA_TRY(flush_pass());
        ZXA_TRY(flush_group());
        if (policy_ == POLICY_FUSED && used_bits_ >= tile_min_bits(prec_))
        {
            wide_u3_ = descs;
            return OK;
        }
        ZXA_TRY(settle());
        return u3_apply(ctx_, dstate_, prec_, active_bits(), descs.data(), descs.size());
    }
    ZXA_TRY(flush_wide());
    const unsigned low = tile_low_bits(prec_);
    const unsigned n = active_bits();
    // Complex128 (X and CX as affine updates): a gate whose target the
    // pending window touches may still join the pass, the window's terms on
    // its target becoming its fan.  A pass with fans keeps its gates in
    // circuit order (tile_staged_fits).  Otherwise a gate the window touches
    // starts a new pass.
    const bool can_fan = prec_ == C128 && !disable_x_;
    // A batch that opens on an empty pass (no window: no fans) and needs more
    // than one pass, with the wide tile on: the passes tile_layer_plan forms,
    // each closed where the plan ends it.  The gates act on distinct qubits
    // of one batch, so any order is the same product.
    std::vector<size_t> breaks;
    if (pass_u3_.empty() && pass_diag_.empty() && descs.size() > tile_upper_slots(prec_) &&
        tune_get(prec_ == C64 ? "tile.wide.c64" : "tile.wide.c128") == 1)
    {
        std::vector<unsigned> ts;
        for (const U3Gate &g : descs)
            ts.push_back(unsigned(__builtin_ctzll(g.delta)));
        std::vector<std::vector<unsigned>> plan;
        if (tile_layer_plan(prec_, n, -1, ts, plan))
        {
            std::vector<U3Gate> ordered;
            for (const std::vector<unsigned> &ps : plan)
            {
                if (!ordered.empty())
                    breaks.push_back(ordered.size());
                for (unsigned t : ps)
                    ordered.push_back(descs[size_t(std::find(ts.begin(), ts.end(), t) - ts.begin())]);
            }
            descs.swap(ordered);
        }
    }
    auto order_with = [&](unsigned t) {
        std::vector<unsigned> v;
        for (const U3Gate &p : pass_u3_)
            v.push_back(unsigned(__builtin_ctzll(p.delta)));
        v.push_back(t);
        return v;
    };
    for (size_t i = 0; i < descs.size(); i++)
    {
        if (std::binary_search(breaks.begin(), breaks.end(), i))
            ZXA_TRY(flush_pass());
        const U3Gate &g0 = descs[i];
        U3Gate g = g0;
        if (g.packed_sel >> 63)
        {
            // beta = 1: the even side is the target-bit-one element
            const double sw[8] = {g0.u[6], g0.u[7], g0.u[4], g0.u[5], g0.u[2], g0.u[3], g0.u[0], g0.u[1]};
            std::memcpy(g.u, sw, sizeof(sw));
            g.packed_sel = g.delta;
        }
        const unsigned t = unsigned(__builtin_ctzll(g.delta));
        int same = -1;
        for (size_t k = 0; k < pass_u3_.size(); k++)
            if (pass_u3_[k].delta == g.delta)
                same = int(k);
        // a later gate's fan on t lies between the two gates on t
        bool hits = (g.delta & pass_support_) != 0;
        if (same >= 0)
            for (size_t k = size_t(same) + 1; k < pass_fan_.size(); k++)
                hits = hits || (pass_fan_[k].on && pass_fan_[k].theta[t] != 0.0);
        GateFan fan;
        if (hits)
        {
            bool ok = can_fan && same < 0 && tile_staged_fits(prec_, n, order_with(t));
            ok = ok && take_fan(t, fan);
            if (!ok)
            {
                ZXA_TRY(flush_pass());
                same = -1;
                fan = GateFan();
            }
        }
        if (same >= 0)
        {
            U3Gate *o = &pass_u3_[size_t(same)];
            const std::complex<double> a[4] = {{g.u[0], g.u[1]}, {g.u[2], g.u[3]}, {g.u[4], g.u[5]}, {g.u[6], g.u[7]}};
            const std::complex<double> b[4] = {{o->u[0], o->u[1]}, {o->u[2], o->u[3]},
                                               {o->u[4], o->u[5]}, {o->u[6], o->u[7]}};
            const std::complex<double> c[4] = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3],
                                               a[2] * b[0] + a[3] * b[2], a[2] * b[1] + a[3] * b[3]};
            for (int i = 0; i < 4; i++)
            {
                o->u[2 * i] = c[i].real();
                o->u[2 * i + 1] = c[i].imag();
            }
            continue;
        }
        const bool high = t >= low;
        // (a pass already wider than the 4096-unit tile takes no low target)
        bool fits = (pass_has_fan_ || fan.on) ? tile_staged_fits(prec_, n, order_with(t))
                                              : unsigned(pass_high_) + (high ? 1u : 0u) <= tile_upper_slots(prec_);
        if (!fits && !pass_has_fan_ && !fan.on &&
            window_is_polynomial(n, pass_diag_.data(), pass_diag_.size()))
        {
            // the wide tile: up to nine targets above a 16-unit run
            std::vector<unsigned> ts = order_with(t);
            fits = tile_wide_fits(prec_, n, ts);
        }
        if (!fits)
        {
            // (a fan was taken only where the staged order fits)
            ZXA_TRY(flush_pass());
            fan = GateFan();
        }
        pass_u3_.push_back(g);
        pass_fan_.push_back(fan);
        pass_has_fan_ = pass_has_fan_ || fan.on;
        if (high)
            pass_high_++;
    }
    return OK;
}

// flush_u3_batch (zxh.cpp:390-428): one descriptor per active accumulator,
// in qubit order; the batch is applied as tile passes.
int Sim::flush_u3(std::vector<Acc> &pend, size_t &count)
{
    if (count == 0)
        return OK;
    const unsigned lb = active_bits();
    std::vector<U3Gate> descs;
    for (size_t q = 0; q < N_; q++)
    {
        Acc &a = pend[q];
        if (!a.active)
            continue;
        U3Gate g;
        if (a.fixed)
        {
            g.packed_sel = a.sel;
            g.delta = a.delta;
        }
        else
        {
            uint64_t d = 0;
            if (!solve(q, &d))
            {
                set_error("flush_u3_batch requires no-expand U3 descriptors");
                return ESTATE;
            }
            g.packed_sel = local_selector(q);
            g.delta = d & low_mask(lb);
        }
        const std::complex<double> u[4] = {a.u00, a.u01, a.u10, a.u11};
        for (int i = 0; i < 4; i++)
        {
            g.u[2 * i] = u[i].real();
            g.u[2 * i + 1] = u[i].imag();
        }
        descs.push_back(g);
    }
    for (Acc &a : pend)
        a = Acc();
    count = 0;
    if (descs.empty())
        return OK;
    return emit_u3(descs);
}

int Sim::execute()
{
    executed_ = false;
    result_count_ = 0;
    const size_t target_M = eager_ ? N_ : required_M();
    ZXA_TRY(resize_reset(target_M));
    gphase_ = {1.0, 0.0};
    std::fill(rows_.begin(), rows_.end(), 0);
    std::fill(b_.begin(), b_.end(), 0);
    m_ = 0;
    if (eager_) // expand_all (zxh.cpp:721-729)
    {
        used_bits_ = N_;
        for (size_t q = 0; q < N_; q++)
            rows_[q] = uint64_t(1) << q;
        m_ = N_;
    }

    pass_u3_.clear();
    pass_fan_.clear();
    pass_diag_.clear();
    pass_support_ = 0;
    pass_high_ = 0;
    pass_has_fan_ = false;
    exp_group_.clear();
    wide_u3_.clear();

    // Scheduling.  The reference policy batches each run of consecutive U3
    // gates and flushes the diagonal window when a run starts (zxh.cpp:440-470).
    // The fused policy keeps a batch open across later gates: pending are
    // the diagonal gates to apply before the batch (pre), the batch (one
    // product per qubit and run, each with the descriptor of its own place in
    // the circuit) and the diagonal gates to apply after it (post), and the
    // order pre, batch, post is a valid reordering of the circuit throughout:
    //  - X and CX as affine updates change only the map;
    //  - a diagonal gate joins pre when it commutes with every member
    //    (selector . delta = 0 over GF(2)), else post;
    //  - a U3 gate joins the batch when it commutes with every gate in post
    //    and every member of another run (alpha . delta' = alpha' . delta = 0),
    //    or continues its own run's product on its qubit.  Otherwise pre and
    //    the members of earlier runs are applied, and post becomes pre (no
    //    diagonal gate arrives inside a run, so the current run's members
    //    follow post in the circuit and commute with it).
    // Each run keeps its own product per qubit, as the reference forms it,
    // so the 1e-5 identity skip sees the same products.
    const bool stream = policy_ == POLICY_FUSED;
    // expansions applied at once (u3_expand_group_apply)
    constexpr size_t kGroupMax = 4;
    std::vector<DiagGate> pre, post;
    std::vector<Acc> pend(N_);
    size_t pend_count = 0;
    std::vector<size_t> owner(N_, 0);
    size_t run = 0;
    bool in_run = false;

    // absorb_u3_gate + left_multiply_u3 (zxh.cpp:144-156, 363-388): the
    // accumulator restarts from the identity once it has come within 1e-5 of it
    auto absorb = [&](size_t q, const std::complex<double> (&u)[4], uint64_t sel, uint64_t delta) {
        Acc &a = pend[q];
        const bool was = a.active;
        if (!was)
        {
            a = Acc();
            a.active = true;
            a.fixed = true;
            a.sel = sel;
            a.delta = delta;
        }
        const std::complex<double> o00 = a.u00, o01 = a.u01, o10 = a.u10, o11 = a.u11;
        a.u00 = u[0] * o00 + u[1] * o10;
        a.u01 = u[0] * o01 + u[1] * o11;
        a.u10 = u[2] * o00 + u[3] * o10;
        a.u11 = u[2] * o01 + u[3] * o11;
        a.active = !(close_val(a.u00, {1.0, 0.0}) && close_val(a.u01, {0.0, 0.0}) && close_val(a.u10, {0.0, 0.0}) &&
                     close_val(a.u11, {1.0, 0.0}));
        if (!was && a.active)
            pend_count++;
        else if (was && !a.active)
            pend_count--;
        owner[q] = run;
    };
    const uint64_t amask = (uint64_t(1) << 63) - 1;
    auto par = [](uint64_t x) { return (__builtin_popcountll(x) & 1) != 0; };
    auto diag_commutes = [&](const DiagGate &d, uint64_t delta) {
        return !par(d.s0 & amask & delta) && (d.s1 == kNoSel || !par(d.s1 & amask & delta));
    };
    auto diag_joins_pre = [&](const DiagGate &d) {
        for (size_t k = 0; k < N_; k++)
            if (pend[k].active && !diag_commutes(d, pend[k].delta))
                return false;
        return true;
    };
    auto push_diag = [&](const DiagGate &d) {
        // a selector that is identically zero (alpha = 0, beta = 0: a qubit not
        // yet expanded) makes the gate's phase zero everywhere
        if (stream && (d.s0 == 0 || (d.s1 != kNoSel && d.s1 == 0)))
            return;
        if (stream && !diag_joins_pre(d))
            post.push_back(d);
        else
            pre.push_back(d);
    };
    // everything pending, in order
    auto flush_all = [&]() -> int {
        ZXA_TRY(flush_diag(pre));
        ZXA_TRY(flush_u3(pend, pend_count));
        return flush_diag(post);
    };
    // pre and the members of earlier runs; post becomes pre
    auto flush_earlier = [&]() -> int {
        std::vector<std::pair<size_t, Acc>> keep;
        for (size_t k = 0; k < N_; k++)
            if (pend[k].active && owner[k] == run)
            {
                keep.push_back({k, pend[k]});
                pend[k] = Acc();
                pend_count--;
            }
        ZXA_TRY(flush_diag(pre));
        ZXA_TRY(flush_u3(pend, pend_count));
        for (const auto &kv : keep)
        {
            pend[kv.first] = kv.second;
            pend_count++;
        }
        pre.swap(post);
        post.clear();
        return OK;
    };

    for (size_t gi = 0; gi < gates_.size(); gi++)
    {
        const GateRec &g = gates_[gi];
        const bool is_u3 = g.type == G_H || g.type == G_U3;
        if (!is_u3 && in_run)
        {
            // end of a run: the reference flushes its batch here
            in_run = false;
            if (!stream)
                ZXA_TRY(flush_u3(pend, pend_count));
        }
        if (is_u3)
        {
            // U3 cluster (zxh.cpp:440-470)
            if (!in_run)
            {
                in_run = true;
                run = gi;
                if (!stream)
                    ZXA_TRY(flush_diag(pre));
            }
            const double th = g.type == G_H ? kPi / 2.0 : g.theta;
            const double la = g.type == G_H ? kPi : g.lambda;
            const double ph = g.type == G_H ? 0.0 : g.phi;
            std::complex<double> u[4];
            u3_matrix(th, la, ph, u);
            uint64_t delta = 0;
            if (!solve(g.q, &delta))
            {
                // flush the batch, then the eager, expanding apply_u3.  While
                // nothing else is pending, up to four expansions in a row wait
                // to be applied at once: one read of the old state and one
                // write of the new instead of a sweep per bit.  Diagonal gates
                // waiting in pre may stay: they act on the old bits only, so
                // they commute with an expansion (both halves of each new pair
                // get the same phase) and are applied after the group
                const bool idle = pend_count == 0 && (stream || pre.empty()) && post.empty() && pass_u3_.empty() &&
                                  pass_diag_.empty() && wide_u3_.empty();
                if (!idle)
                {
                    ZXA_TRY(flush_all());
                    ZXA_TRY(flush_pass());
                }
                const unsigned before = active_bits();
                int rc = OK;
                const uint64_t d = solve_expand(g.q, &rc);
                if (rc)
                    return rc;
                U3Gate eg;
                eg.packed_sel = local_selector(g.q);
                eg.delta = d & low_mask(active_bits());
                for (int i = 0; i < 4; i++)
                {
                    eg.u[2 * i] = u[i].real();
                    eg.u[2 * i + 1] = u[i].imag();
                }
                // the new bit is the top one and its half still zero
                const bool top = eg.delta == uint64_t(1) << (active_bits() - 1);
                if (stream && top && !exp_group_.empty() &&
                    (exp_group_.size() == kGroupMax || exp_base_ + exp_group_.size() != before))
                    ZXA_TRY(flush_group());
                if (stream && top)
                {
                    if (exp_group_.empty())
                        exp_base_ = before;
                    exp_group_.push_back(eg);
                    continue;
                }
                ZXA_TRY(flush_group());
                ZXA_TRY(settle());
                if (top)
                    ZXA_TRY(u3_expand_apply(ctx_, dstate_, prec_, active_bits(), eg));
                else
                    ZXA_TRY(u3_pair_apply(ctx_, dstate_, prec_, active_bits(), eg));
                continue;
            }
            delta &= low_mask(active_bits());
            const uint64_t sel = local_selector(g.q);
            if (stream)
            {
                const uint64_t alpha = sel & amask;
                bool ok = !(pend[g.q].active && owner[g.q] != run);
                for (const DiagGate &d : post)
                    ok = ok && diag_commutes(d, delta);
                for (size_t k = 0; k < N_ && ok; k++)
                    if (pend[k].active && owner[k] != run)
                        ok = !par(alpha & pend[k].delta) && !par(pend[k].sel & amask & delta);
                if (!ok)
                    ZXA_TRY(flush_earlier());
            }
            // absorb_u3_gate + left_multiply_u3 (zxh.cpp:144-156, 363-388)
            absorb(g.q, u, sel, delta);
            continue;
        }
        switch (g.type)
        {
        case G_BARRIER:
            ZXA_TRY(flush_all());
            break;
        case G_Z:
        case G_RZ: {
            const double th = g.type == G_Z ? kPi : g.theta;
            push_diag({local_selector(g.q), kNoSel, th});
            gphase_ *= phase(-0.5 * th);
            break;
        }
        case G_P:
            push_diag({local_selector(g.q), kNoSel, g.theta});
            break;
        case G_X:
            if (disable_x_)
            {
                ZXA_TRY(flush_all());
                ZXA_TRY(flush_pass());
                int rc = OK;
                const uint64_t d = solve_expand(g.q, &rc);
                if (rc)
                    return rc;
                ZXA_TRY(flush_group());
                ZXA_TRY(settle());
                ZXA_TRY(x_apply(ctx_, dstate_, prec_, active_bits(), d & low_mask(active_bits()),
                                local_selector(g.q)));
            }
            else
                b_[g.q] ^= 1u;
            break;
        case G_CX:
            if (disable_x_)
            {
                ZXA_TRY(flush_all());
                ZXA_TRY(flush_pass());
                int rc = OK;
                const uint64_t d = solve_expand(g.q, &rc);
                if (rc)
                    return rc;
                ZXA_TRY(flush_group());
                ZXA_TRY(settle());
                ZXA_TRY(cx_apply(ctx_, dstate_, prec_, active_bits(), d & low_mask(active_bits()),
                                 local_selector(g.cq), local_selector(g.q)));
            }
            else
            {
                rows_[g.q] ^= rows_[g.cq];
                b_[g.q] = uint8_t(b_[g.q] != b_[g.cq]);
            }
            break;
        case G_CP: {
            const uint64_t scq = local_selector(g.cq), sq = local_selector(g.q);
            if (stream && (scq == 0 || sq == 0))
                break; // phase zero everywhere (see push_diag)
            // the reference applies a CP whose selectors mix the low eight
            // and higher address bits in a kernel of its own; the fused
            // policy keeps it in the diagonal window (a table epilogue or a
            // general sweep takes any selectors)
            if (!stream && (sel_class(scq) == 2 || sel_class(sq) == 2))
            {
                ZXA_TRY(flush_all());
                ZXA_TRY(flush_pass());
                ZXA_TRY(flush_group());
                ZXA_TRY(settle());
                ZXA_TRY(cp_apply(ctx_, dstate_, prec_, active_bits(), scq, sq, g.theta));
            }
            else
                push_diag({scq, sq, g.theta});
            break;
        }
        default:
            break;
        }
    }
    ZXA_TRY(flush_all());
    ZXA_TRY(flush_pass());
    ZXA_TRY(flush_group());
    ZXA_TRY(settle());
    ZXA_HIP(hipStreamSynchronize(ctx_.stream));
    executed_ = true;
    return OK;
}

// measure (measure_common.cpp:73-92) with the GPU sampler.
int Sim::measure(size_t cnt)
{
    if (!executed_)
    {
        set_error("measure requires execute() to be called first");
        return ESTATE;
    }
    if (cnt == 0)
    {
        result_count_ = 0;
        return OK;
    }
    const unsigned m = active_bits();
    const unsigned I = unsigned(std::min<size_t>(M_ <= seg_min_i_ ? M_ : seg_min_i_, m));
    result_count_ = 0;
    // the probability tree builds on the GPU while the host draws the
    // uniforms and their factors 1 - u^(1/rest), which do not depend on the
    // total; the recurrence runs once the total is back
    ZXA_TRY(sample_launch(ctx_, dstate_, prec_, m, I));
    const uint64_t base = use_seed_ ? seed_ : uint64_t(std::random_device{}());
    const uint64_t stream_seed = mix_seed(base, 0);
    ZXA_TRY(thr_pin_.ensure(cnt * sizeof(double)));
    double *thr = static_cast<double *>(thr_pin_.ptr);
    threshold_draws(stream_seed, cnt, thr);
    threshold_factors(thr, cnt);
    double total = 0.0;
    ZXA_TRY(sample_total(ctx_, &total));
    if (!(total > 0.0))
    {
        set_error("state has zero total probability");
        return EZEROPROB;
    }
    threshold_apply(thr, cnt, 0.0, total);
    ZXA_TRY(real_pin_.ensure(cnt * sizeof(uint64_t)));
    uint64_t *real = static_cast<uint64_t *>(real_pin_.ptr);
    // virtual result words (pack_virtual_results_host,
    // measure_result_pack.h:24-44): under the identity map (plain method)
    // they are the real addresses; otherwise the device packs them after the
    // search
    const size_t wc = (N_ + 63) / 64;
    bool identity = (b_.end() == std::find(b_.begin(), b_.end(), uint8_t(1))) && N_ <= 64;
    for (size_t q = 0; q < N_ && identity; q++)
        identity = rows_[q] == (uint64_t(1) << q);
    packed_is_real_ = identity;
    if (identity)
    {
        ZXA_TRY(sample_search(ctx_, dstate_, prec_, thr, cnt, real));
    }
    else
    {
        ZXA_TRY(packed_pin_.ensure(cnt * wc * sizeof(uint64_t)));
        PackSpec ps;
        ps.nq = unsigned(N_);
        ps.rows = rows_.data();
        ps.b = b_.data();
        ps.out = static_cast<uint64_t *>(packed_pin_.ptr);
        ZXA_TRY(sample_search(ctx_, dstate_, prec_, thr, cnt, real, &ps));
    }
    result_count_ = cnt;
    return OK;
}

int Sim::get_results(uint64_t *packed, size_t cnt) const
{
    if (cnt > result_count_)
    {
        set_error("get_results count exceeds cached result count");
        return EINVAL_;
    }
    const size_t wc = (N_ + 63) / 64;
    if (cnt == 0)
        return OK;
    std::memcpy(packed, packed_is_real_ ? real_pin_.ptr : packed_pin_.ptr, cnt * wc * sizeof(uint64_t));
    return OK;
}

int Sim::get_real_results(uint64_t *real, size_t cnt) const
{
    if (cnt > result_count_)
    {
        set_error("get_results count exceeds cached result count");
        return EINVAL_;
    }
    if (cnt)
        std::memcpy(real, real_pin_.ptr, cnt * sizeof(uint64_t));
    return OK;
}

int Sim::get_phys_state(void *out, size_t out_bytes, size_t *m_out) const
{
    if (!executed_)
    {
        set_error("get_state requires execute() to be called first");
        return ESTATE;
    }
    const size_t bytes = (size_t(1) << used_bits_) * elem_bytes(prec_);
    if (out_bytes < bytes)
    {
        set_error("output buffer too small");
        return EINVAL_;
    }
    ZXA_HIP(hipMemcpy(out, dstate_, bytes, hipMemcpyDeviceToHost));
    *m_out = used_bits_;
    return OK;
}

int Sim::set_phys_state(const void *in, size_t in_bytes)
{
    if (!executed_)
    {
        set_error("set_phys_state requires execute() to be called first");
        return ESTATE;
    }
    const size_t bytes = (size_t(1) << used_bits_) * elem_bytes(prec_);
    if (in_bytes != bytes)
    {
        set_error("input size does not match the physical state");
        return EINVAL_;
    }
    ZXA_HIP(hipMemcpy(dstate_, in, bytes, hipMemcpyHostToDevice));
    result_count_ = 0;
    return OK;
}

int Sim::device_ptr(void **p, size_t *m) const
{
    if (!executed_)
    {
        set_error("no state before execute()");
        return ESTATE;
    }
    *p = dstate_;
    *m = used_bits_;
    return OK;
}

// get_state (zxh.cpp:579-636): virt = A*real ^ b, stored at the bit-reversed
// index, times the global phase.
int Sim::get_state(double *out, size_t out_len) const
{
    if (!executed_)
    {
        set_error("get_state requires execute() to be called first");
        return ESTATE;
    }
    if (N_ >= 40)
    {
        set_error("get_state full state export requires fewer than 40 qubits");
        return EINVAL_;
    }
    const size_t vsize = size_t(1) << N_;
    if (out_len < 2 * vsize)
    {
        set_error("output buffer too small");
        return EINVAL_;
    }
    const size_t real_size = size_t(1) << m_;
    std::vector<std::complex<double>> real_state(real_size);
    if (prec_ == C128)
    {
        ZXA_HIP(hipMemcpy(real_state.data(), dstate_, real_size * 16, hipMemcpyDeviceToHost));
    }
    else
    {
        std::vector<std::complex<float>> tmp(real_size);
        ZXA_HIP(hipMemcpy(tmp.data(), dstate_, real_size * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < real_size; i++)
            real_state[i] = std::complex<double>(tmp[i].real(), tmp[i].imag());
    }
    std::fill(out, out + 2 * vsize, 0.0);
    for (size_t real = 0; real < real_size; real++)
    {
        uint64_t virt = 0;
        for (size_t q = 0; q < N_; q++)
            virt |= uint64_t((__builtin_popcountll(real & rows_[q]) & 1) ^ b_[q]) << q;
        uint64_t rev = 0;
        for (size_t i = 0; i < N_; i++)
            rev = (rev << 1) | ((virt >> i) & 1u);
        const std::complex<double> v = real_state[real] * gphase_;
        out[2 * rev] = v.real();
        out[2 * rev + 1] = v.imag();
    }
    return OK;
}

} // namespace zxa
