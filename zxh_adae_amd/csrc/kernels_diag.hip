// kernels_diag.hip -- the diagonal-window sweep (P / CP / Rz / Z batches).
//
// Replaces diag_block_kernel_cuda and the CPU apply_diag_window (reference
// src/zxh/src/kernel/kernel_cuda.cu:505-614, kernel_general.cpp:72-113): every
// amplitude x is multiplied by exp(i*theta(x)) with
//     theta(x) = sum_g theta_g * [S0_g(x)] * [S1_g(x)],
//     S(x) = parity(x & alpha) ^ beta          (S1 == 1 for a P gate),
// in one read-modify-write sweep over the state.
//
// Fast path (every selector has at most one alpha bit, the plain method and
// most ZXH windows): theta(x) is then a polynomial of degree <= 2 in the
// address bits,
//     theta(x) = K + sum_j L_j x_j + sum_{j<k} Q_jk x_j x_k .
// The address splits into lane bits a (16 contiguous bytes per lane, 64
// lanes), row bits w (64 rows of 1 KiB per chunk) and chunk bits h.  Then
//     theta = lane(a; h) + rowlo(a; w_lo) + row(w; h)
// where lane() is per lane and chunk, row() is uniform per row of the chunk,
// and rowlo() is linear in the four row bits a wave iterates over.  Each
// factor becomes a unit complex number computed once per lane or row, so an
// amplitude costs three complex products (no sincos per element), and the
// sweep stays HBM-bound at any batch size.
//
// General path (multi-bit selectors after CX transport, or states narrower
// than one chunk): per-amplitude evaluation of every gate, like the reference.
#include "dev.h"
#include "poly.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace zxa
{
namespace
{

constexpr int kThreads = 256;
constexpr int kRowBits = 6;   // 64 rows per chunk
constexpr int kRowsPerWave = 16;
constexpr int kCoefStride = 16; // doubles per chunk in the coefficient array

struct cd
{
    double x, y;
};
struct cf
{
    float x, y;
};

__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf cmulf(cf a, cf b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cd cis(double a)
{
    double s, c;
    sincos(a, &s, &c);
    return {c, s};
}
__device__ __forceinline__ cf tof(cd a)
{
    return {static_cast<float>(a.x), static_cast<float>(a.y)};
}
// exp(i a) for a complex64 state: the angle is summed and reduced modulo
// 2 pi in double precision, the sine and cosine are taken in single precision
// (a few ulps of float, the precision of the amplitudes they multiply; the
// double-precision sincos made the complex64 sweep VALU-bound)
__device__ __forceinline__ cf cis_f(double a)
{
    const double r = a - 6.283185307179586 * rint(a * 0.15915494309189535);
    float s, c;
    sincosf(static_cast<float>(r), &s, &c);
    return {c, s};
}
__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffu), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                static_cast<unsigned>(lo));
}

template <bool kC64>
struct Geo
{
    static constexpr int EPL = kC64 ? 2 : 1;     // amplitudes per lane (16 bytes)
    static constexpr int LB = 6 + (kC64 ? 1 : 0); // lane-element bits
    static constexpr int CB = LB + kRowBits;      // chunk bits
    static constexpr int NCO = LB + kRowBits;     // per-chunk linear coefficients
};

// Per-chunk coefficients: c_j(h) = sum_{k in h} Q_jk h_k for every bit j of
// a and w, and C(h) = K + sum_{k in h} h_k (L_k + sum_{l>k} Q_kl h_l).
__global__ __launch_bounds__(kThreads) void diag_chunk_coef_kernel(const double *__restrict__ cq,
                                                                   const double *__restrict__ hL,
                                                                   const double *__restrict__ hQ, double K, int nco,
                                                                   int nh, uint64_t nchunks, double *__restrict__ out)
{
    const uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (c >= nchunks)
        return;
    double *o = out + c * kCoefStride;
    for (int j = 0; j < nco; j++)
    {
        double s = 0.0;
        for (int i = 0; i < nh; i++)
            if ((c >> i) & 1u)
                s += cq[j * nh + i];
        o[j] = s;
    }
    double C = K;
    for (int i = 0; i < nh; i++)
    {
        if (!((c >> i) & 1u))
            continue;
        double t = hL[i];
        for (int l = i + 1; l < nh; l++)
            if ((c >> l) & 1u)
                t += hQ[i * nh + l];
        C += t;
    }
    o[nco] = C;
}

// tabs: lane_tab[2^LB] | dco[2^LB * 6] | row_tab[64]
template <bool kC64>
__global__ __launch_bounds__(kThreads) void diag_poly_kernel(void *__restrict__ state, const double *__restrict__ tabs,
                                                             const double *__restrict__ coef, uint64_t nchunks)
{
    using G = Geo<kC64>;
    constexpr int EPL = G::EPL;
    constexpr int LB = G::LB;
    const double *lane_tab = tabs;
    const double *dco = tabs + (1 << LB);
    const double *row_tab = dco + (1 << LB) * kRowBits;

    const int lane = threadIdx.x & 63;
    const int v = threadIdx.x >> 6;

    // chunk-independent per-lane factors
    double aw[EPL];
    cd glo[EPL][4], ghi[EPL][4];
#pragma unroll
    for (int e = 0; e < EPL; e++)
    {
        const int xa = lane * EPL + e;
        const double *d = dco + xa * kRowBits;
        aw[e] = lane_tab[xa] + ((v & 1) ? d[4] : 0.0) + ((v & 2) ? d[5] : 0.0);
        glo[e][0] = {1.0, 0.0};
        glo[e][1] = cis(d[0]);
        glo[e][2] = cis(d[1]);
        glo[e][3] = cis(d[0] + d[1]);
        ghi[e][0] = {1.0, 0.0};
        ghi[e][1] = cis(d[2]);
        ghi[e][2] = cis(d[3]);
        ghi[e][3] = cis(d[2] + d[3]);
    }
    const int xw_own = v * kRowsPerWave + (lane & (kRowsPerWave - 1));
    const double brow = row_tab[xw_own];

    for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x)
    {
        const size_t base = static_cast<size_t>(chunk) << G::CB;
        const double *cc = coef + chunk * kCoefStride;

        if constexpr (!kC64)
        {
            double2 *p = static_cast<double2 *>(state);
            double2 data[kRowsPerWave];
#pragma unroll
            for (int t = 0; t < kRowsPerWave; t++)
                data[t] = p[base + (static_cast<size_t>(v * kRowsPerWave + t) << LB) + lane];

            double ang = aw[0];
#pragma unroll
            for (int j = 0; j < LB; j++)
                if ((lane >> j) & 1)
                    ang += cc[j];
            const cd E = cis(ang);
            double rang = brow + cc[G::NCO];
#pragma unroll
            for (int k = 0; k < kRowBits; k++)
                if ((xw_own >> k) & 1)
                    rang += cc[LB + k];
            const cd R = cis(rang);
            cd q[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                q[u] = cmul(E, ghi[0][u]);
#pragma unroll
            for (int t = 0; t < kRowsPerWave; t++)
            {
                const cd rt = {readlane_d(R.x, t), readlane_d(R.y, t)};
                const cd ph = cmul(cmul(q[t >> 2], glo[0][t & 3]), rt);
                const cd a = cmul({data[t].x, data[t].y}, ph);
                data[t].x = a.x;
                data[t].y = a.y;
            }
#pragma unroll
            for (int t = 0; t < kRowsPerWave; t++)
                p[base + (static_cast<size_t>(v * kRowsPerWave + t) << LB) + lane] = data[t];
        }
        else
        {
            float4 *p = static_cast<float4 *>(state);
            const size_t base4 = base >> 1;
            float4 data[kRowsPerWave];
#pragma unroll
            for (int t = 0; t < kRowsPerWave; t++)
                data[t] = p[base4 + (static_cast<size_t>(v * kRowsPerWave + t) << (LB - 1)) + lane];

            cf q[EPL][4];
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                const int xa = lane * EPL + e;
                double ang = aw[e];
#pragma unroll
                for (int j = 0; j < LB; j++)
                    if ((xa >> j) & 1)
                        ang += cc[j];
                const cd E = cis(ang);
#pragma unroll
                for (int u = 0; u < 4; u++)
                    q[e][u] = tof(cmul(E, ghi[e][u]));
            }
            double rang = brow + cc[G::NCO];
#pragma unroll
            for (int k = 0; k < kRowBits; k++)
                if ((xw_own >> k) & 1)
                    rang += cc[LB + k];
            const cd R = cis(rang);
            const float rx = static_cast<float>(R.x), ry = static_cast<float>(R.y);
            cf gl[EPL][4];
#pragma unroll
            for (int e = 0; e < EPL; e++)
#pragma unroll
                for (int u = 0; u < 4; u++)
                    gl[e][u] = tof(glo[e][u]);
#pragma unroll
            for (int t = 0; t < kRowsPerWave; t++)
            {
                const cf rt = {__int_as_float(__builtin_amdgcn_readlane(__float_as_int(rx), t)),
                               __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ry), t))};
                const cf ph0 = cmulf(cmulf(q[0][t >> 2], gl[0][t & 3]), rt);
                const cf ph1 = cmulf(cmulf(q[1][t >> 2], gl[1][t & 3]), rt);
                const cf a0 = cmulf({data[t].x, data[t].y}, ph0);
                const cf a1 = cmulf({data[t].z, data[t].w}, ph1);
                data[t] = make_float4(a0.x, a0.y, a1.x, a1.y);
            }
#pragma unroll
            for (int t = 0; t < kRowsPerWave; t++)
                p[base4 + (static_cast<size_t>(v * kRowsPerWave + t) << (LB - 1)) + lane] = data[t];
        }
    }
}

// Flat-grid variant: one workgroup per chunk of 256*EPT sixteen-byte units
// (4 waves x EPT rows of 1 KiB), launched as one workgroup per chunk so that
// the hardware dispatcher walks the state in address order.  Same phase
// factorisation: lane phase (per lane and chunk), row phase (per row, one
// sincos by lane t, broadcast), and a chunk-independent lane x row factor
// read from a small table (L1/L2-resident).
// tabs: av[4][2^LB] | brow[4*EPT] | g[2^LB][EPT] (complex, re/im pairs)
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

// 16-byte loads/stores, optionally with the nontemporal hint.
template <bool kNT>
__device__ __forceinline__ double2 ld16(const double2 *p)
{
    const dvec2 v = kNT ? __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p))
                        : *reinterpret_cast<const dvec2 *>(p);
    return make_double2(v.x, v.y);
}
template <bool kNT>
__device__ __forceinline__ void st16(double2 a, double2 *p)
{
    dvec2 v;
    v.x = a.x;
    v.y = a.y;
    if (kNT)
        __builtin_nontemporal_store(v, reinterpret_cast<dvec2 *>(p));
    else
        *reinterpret_cast<dvec2 *>(p) = v;
}
template <bool kNT>
__device__ __forceinline__ float4 ld16(const float4 *p)
{
    const fvec4 v = kNT ? __builtin_nontemporal_load(reinterpret_cast<const fvec4 *>(p))
                        : *reinterpret_cast<const fvec4 *>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}
template <bool kNT>
__device__ __forceinline__ void st16(float4 a, float4 *p)
{
    fvec4 v;
    v.x = a.x;
    v.y = a.y;
    v.z = a.z;
    v.w = a.w;
    if (kNT)
        __builtin_nontemporal_store(v, reinterpret_cast<fvec4 *>(p));
    else
        *reinterpret_cast<fvec4 *>(p) = v;
}

template <bool kC64, int kEPT, bool kNT>
__global__ __launch_bounds__(kThreads) void diag_flat_kernel(void *__restrict__ state, const double *__restrict__ tabs,
                                                             const double *__restrict__ coef)
{
    constexpr int EPL = kC64 ? 2 : 1;
    constexpr int LB = 6 + (kC64 ? 1 : 0);
    constexpr int RB = 2 + (kEPT == 1 ? 0 : kEPT == 2 ? 1 : kEPT == 4 ? 2 : 3);
    constexpr int NCO = LB + RB;
    const double *av = tabs;
    const double *brow = av + 4 * (1 << LB);
    const double *gt = brow + 4 * kEPT;

    const int lane = threadIdx.x & 63;
    const int v = threadIdx.x >> 6;
    const uint64_t chunk = blockIdx.x;
    const size_t base = static_cast<size_t>(chunk) << (LB + RB);
    const double *cc = coef + chunk * kCoefStride;

    // loads first: they do not depend on the phases
    float4 d4[kEPT];
    double2 d2[kEPT];
#pragma unroll
    for (int t = 0; t < kEPT; t++)
    {
        const size_t row = static_cast<size_t>(v * kEPT + t);
        if constexpr (kC64)
            d4[t] = ld16<kNT>(static_cast<const float4 *>(state) + ((base + (row << LB)) >> 1) + lane);
        else
            d2[t] = ld16<kNT>(static_cast<const double2 *>(state) + base + (row << LB) + lane);
    }
    // row phase of row v*EPT + t, computed by lane t of the wave
    const int xw_own = v * kEPT + (lane & (kEPT - 1));
    double rang = brow[xw_own] + cc[NCO];
#pragma unroll
    for (int k = 0; k < RB; k++)
        if ((xw_own >> k) & 1)
            rang += cc[LB + k];
    if constexpr (kC64)
    {
        const cf R = cis_f(rang);
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            const int xa = lane * EPL + e;
            double ang = av[v * (1 << LB) + xa];
#pragma unroll
            for (int j = 0; j < LB; j++)
                if ((xa >> j) & 1)
                    ang += cc[j];
            const cf E = cis_f(ang);
#pragma unroll
            for (int t = 0; t < kEPT; t++)
            {
                const double *g = gt + 2 * (xa * kEPT + t);
                const cf rt = {readlane_f(R.x, t), readlane_f(R.y, t)};
                const cf ph = cmulf(cmulf(E, {static_cast<float>(g[0]), static_cast<float>(g[1])}), rt);
                float &re = e == 0 ? d4[t].x : d4[t].z;
                float &im = e == 0 ? d4[t].y : d4[t].w;
                const float a = re, b = im;
                re = a * ph.x - b * ph.y;
                im = a * ph.y + b * ph.x;
            }
        }
    }
    else
    {
        const cd R = cis(rang);
        const int xa = lane;
        double ang = av[v * (1 << LB) + xa];
#pragma unroll
        for (int j = 0; j < LB; j++)
            if ((xa >> j) & 1)
                ang += cc[j];
        const cd E = cis(ang);
#pragma unroll
        for (int t = 0; t < kEPT; t++)
        {
            const double *g = gt + 2 * (xa * kEPT + t);
            const cd rt = {readlane_d(R.x, t), readlane_d(R.y, t)};
            const cd ph = cmul(cmul(E, {g[0], g[1]}), rt);
            const double a = d2[t].x, b = d2[t].y;
            d2[t].x = a * ph.x - b * ph.y;
            d2[t].y = a * ph.y + b * ph.x;
        }
    }
#pragma unroll
    for (int t = 0; t < kEPT; t++)
    {
        const size_t row = static_cast<size_t>(v * kEPT + t);
        if constexpr (kC64)
            st16<kNT>(d4[t], static_cast<float4 *>(state) + ((base + (row << LB)) >> 1) + lane);
        else
            st16<kNT>(d2[t], static_cast<double2 *>(state) + base + (row << LB) + lane);
    }
}

__device__ __forceinline__ bool sel_eval(uint64_t packed, uint64_t x)
{
    const uint64_t alpha = packed & ((uint64_t(1) << 63) - 1);
    return ((__popcll(x & alpha) & 1u) ^ static_cast<unsigned>(packed >> 63)) != 0u;
}

// General path: theta(x) accumulated over every gate of the window.
template <bool kC64>
__global__ __launch_bounds__(kThreads) void diag_general_kernel(void *__restrict__ state,
                                                                const DiagGate *__restrict__ gates, int ngates,
                                                                uint64_t nelems)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; x < nelems; x += stride)
    {
        double th = 0.0;
        for (int g = 0; g < ngates; g++)
        {
            const DiagGate gg = gates[g];
            if (sel_eval(gg.s0, x) && (gg.s1 == kNoSel || sel_eval(gg.s1, x)))
                th += gg.theta;
        }
        if (th == 0.0)
            continue;
        const cd ph = cis(th);
        if constexpr (!kC64)
        {
            double2 *p = static_cast<double2 *>(state);
            const double2 a = p[x];
            const cd r = cmul({a.x, a.y}, ph);
            p[x] = make_double2(r.x, r.y);
        }
        else
        {
            float2 *p = static_cast<float2 *>(state);
            const float2 a = p[x];
            const cf r = cmulf({a.x, a.y}, tof(ph));
            p[x] = make_float2(r.x, r.y);
        }
    }
}

template <typename K>
int persistent_grid(Ctx &ctx, K kernel, uint64_t work_items)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 2;
    const uint64_t cap = static_cast<uint64_t>(per_cu) * static_cast<uint64_t>(ctx.num_cus > 0 ? ctx.num_cus : 256);
    return static_cast<int>(work_items < cap ? work_items : cap);
}

template <bool kC64, int kEPT>
int launch_flat(Ctx &ctx, void *state, const double *tabs, const double *coef, uint64_t nchunks, bool nt)
{
    if (nt)
        hipLaunchKernelGGL((diag_flat_kernel<kC64, kEPT, true>), dim3(static_cast<unsigned>(nchunks)), dim3(kThreads),
                           0, ctx.stream, state, tabs, coef);
    else
        hipLaunchKernelGGL((diag_flat_kernel<kC64, kEPT, false>), dim3(static_cast<unsigned>(nchunks)),
                           dim3(kThreads), 0, ctx.stream, state, tabs, coef);
    return hip_check(hipGetLastError(), "diag_flat_kernel launch");
}

// Flat-grid diagonal sweep with EPT sixteen-byte units per thread.
int diag_flat(Ctx &ctx, void *state, int prec, unsigned n, const DiagGate *gates, size_t ngates, int ept, bool nt,
              double *ws)
{
    const bool c64 = prec == C64;
    const int LB = c64 ? 7 : 6;
    const int RB = 2 + (ept == 1 ? 0 : ept == 2 ? 1 : 2);
    const int NCO = LB + RB;
    const int nh = static_cast<int>(n) - NCO;
    Poly P;
    window_polynomial(n, gates, ngates, P);
    const int nlane = 1 << LB;
    std::vector<double> blk;
    // av[v][xa]: lane part incl. the two wave-select row bits (row bits RB-2, RB-1)
    const int tb = RB - 2;
    for (int vv = 0; vv < 4; vv++)
        for (int xa = 0; xa < nlane; xa++)
        {
            double s = 0.0;
            for (int j = 0; j < LB; j++)
            {
                if (!((xa >> j) & 1))
                    continue;
                s += P.L[j];
                for (int k = j + 1; k < LB; k++)
                    if ((xa >> k) & 1)
                        s += P.q(j, k);
                for (int w = 0; w < 2; w++)
                    if ((vv >> w) & 1)
                        s += P.q(j, LB + tb + w);
            }
            blk.push_back(s);
        }
    // brow[xw]: row part within the chunk
    for (int xw = 0; xw < (1 << RB); xw++)
    {
        double s = 0.0;
        for (int k = 0; k < RB; k++)
        {
            if (!((xw >> k) & 1))
                continue;
            s += P.L[LB + k];
            for (int l = k + 1; l < RB; l++)
                if ((xw >> l) & 1)
                    s += P.q(LB + k, LB + l);
        }
        blk.push_back(s);
    }
    // g[xa][t]: exp(i sum_{k < tb} t_k Q(xa bits, row bit k)), chunk-independent
    for (int xa = 0; xa < nlane; xa++)
        for (int t = 0; t < ept; t++)
        {
            double a = 0.0;
            for (int j = 0; j < LB; j++)
                if ((xa >> j) & 1)
                    for (int k = 0; k < tb; k++)
                        if ((t >> k) & 1)
                            a += P.q(j, LB + k);
            blk.push_back(std::cos(a));
            blk.push_back(std::sin(a));
        }
    void *dblk = nullptr;
    ZXA_TRY(ctx.ring.stage(blk.data(), blk.size() * sizeof(double), ctx.stream, &dblk));
    std::vector<int> inner(NCO), outer(nh);
    for (int j = 0; j < NCO; j++)
        inner[j] = j;
    for (int i = 0; i < nh; i++)
        outer[i] = NCO + i;
    const uint64_t nchunks = uint64_t(1) << nh;
    double *coef = ws;
    if (!coef)
    {
        ZXA_TRY(ctx.scratch_a.ensure(nchunks * kCoefStride * sizeof(double)));
        coef = static_cast<double *>(ctx.scratch_a.ptr);
    }
    ZXA_TRY(launch_block_coef(ctx, P, inner.data(), NCO, outer.data(), nh, coef));
    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    const double *tabs = static_cast<const double *>(dblk);
    ZXA_TRY(ctx.tic());
    int rc;
    if (c64)
        rc = ept == 1 ? launch_flat<true, 1>(ctx, state, tabs, coef, nchunks, nt)
             : ept == 2 ? launch_flat<true, 2>(ctx, state, tabs, coef, nchunks, nt)
                        : launch_flat<true, 4>(ctx, state, tabs, coef, nchunks, nt);
    else
        rc = ept == 1 ? launch_flat<false, 1>(ctx, state, tabs, coef, nchunks, nt)
             : ept == 2 ? launch_flat<false, 2>(ctx, state, tabs, coef, nchunks, nt)
                        : launch_flat<false, 4>(ctx, state, tabs, coef, nchunks, nt);
    ZXA_TRY(rc);
    ZXA_TRY(ctx.toc(F_DIAG, bytes));
    return ctx.ring.done(ctx.stream);
}

int diag_fast(Ctx &ctx, void *state, int prec, unsigned n, const DiagGate *gates, size_t ngates, double *ws)
{
    const bool c64 = prec == C64;
    const int LB = c64 ? Geo<true>::LB : Geo<false>::LB;
    const int NCO = LB + kRowBits;
    const int nh = static_cast<int>(n) - NCO;
    Poly P;
    window_polynomial(n, gates, ngates, P);

    const int nlane = 1 << LB;
    std::vector<double> blk;
    blk.reserve(nlane * (1 + kRowBits) + 64);
    // lane_tab
    for (int xa = 0; xa < nlane; xa++)
    {
        double s = 0.0;
        for (int j = 0; j < LB; j++)
        {
            if (!((xa >> j) & 1))
                continue;
            s += P.L[j];
            for (int k = j + 1; k < LB; k++)
                if ((xa >> k) & 1)
                    s += P.q(j, k);
        }
        blk.push_back(s);
    }
    // dco
    for (int xa = 0; xa < nlane; xa++)
        for (int k = 0; k < kRowBits; k++)
        {
            double s = 0.0;
            for (int j = 0; j < LB; j++)
                if ((xa >> j) & 1)
                    s += P.q(j, LB + k);
            blk.push_back(s);
        }
    // row_tab
    for (int xw = 0; xw < (1 << kRowBits); xw++)
    {
        double s = 0.0;
        for (int k = 0; k < kRowBits; k++)
        {
            if (!((xw >> k) & 1))
                continue;
            s += P.L[LB + k];
            for (int l = k + 1; l < kRowBits; l++)
                if ((xw >> l) & 1)
                    s += P.q(LB + k, LB + l);
        }
        blk.push_back(s);
    }
    void *dblk = nullptr;
    ZXA_TRY(ctx.ring.stage(blk.data(), blk.size() * sizeof(double), ctx.stream, &dblk));

    std::vector<int> inner(NCO), outer(nh);
    for (int j = 0; j < NCO; j++)
        inner[j] = j;
    for (int i = 0; i < nh; i++)
        outer[i] = NCO + i;
    const uint64_t nchunks = uint64_t(1) << nh;
    double *coef = ws;
    if (!coef)
    {
        ZXA_TRY(ctx.scratch_a.ensure(nchunks * kCoefStride * sizeof(double)));
        coef = static_cast<double *>(ctx.scratch_a.ptr);
    }
    ZXA_TRY(launch_block_coef(ctx, P, inner.data(), NCO, outer.data(), nh, coef));

    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    if (c64)
    {
        const int grid = persistent_grid(ctx, diag_poly_kernel<true>, nchunks);
        hipLaunchKernelGGL(diag_poly_kernel<true>, dim3(grid), dim3(kThreads), 0, ctx.stream, state,
                           static_cast<const double *>(dblk), static_cast<const double *>(coef), nchunks);
    }
    else
    {
        const int grid = persistent_grid(ctx, diag_poly_kernel<false>, nchunks);
        hipLaunchKernelGGL(diag_poly_kernel<false>, dim3(grid), dim3(kThreads), 0, ctx.stream, state,
                           static_cast<const double *>(dblk), static_cast<const double *>(coef), nchunks);
    }
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_DIAG, bytes));
    return ctx.ring.done(ctx.stream);
}

int diag_general(Ctx &ctx, void *state, int prec, unsigned n, const DiagGate *gates, size_t ngates)
{
    void *dg = nullptr;
    ZXA_TRY(ctx.ring.stage(gates, ngates * sizeof(DiagGate), ctx.stream, &dg));
    const uint64_t nelems = uint64_t(1) << n;
    const uint64_t blocks = (nelems + kThreads - 1) / kThreads;
    const unsigned grid = static_cast<unsigned>(blocks < 65536 ? blocks : 65536);
    const uint64_t bytes = 2ull * nelems * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(diag_general_kernel<true>, dim3(grid), dim3(kThreads), 0, ctx.stream, state,
                           static_cast<const DiagGate *>(dg), static_cast<int>(ngates), nelems);
    else
        hipLaunchKernelGGL(diag_general_kernel<false>, dim3(grid), dim3(kThreads), 0, ctx.stream, state,
                           static_cast<const DiagGate *>(dg), static_cast<int>(ngates), nelems);
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_DIAG, bytes));
    return ctx.ring.done(ctx.stream);
}

} // namespace

// Degree <= 2 in the address bits: a two-selector gate whose selectors have
// at most one bit each, or a one-selector gate (P, Rz, Z) whose selector has
// at most two bits ([x_u ^ x_v] = x_u + x_v - 2 x_u x_v; the Rz of an
// Rzz lowered as CX Rz CX with X and CX as affine updates is one).
bool window_is_polynomial(unsigned n, const DiagGate *gates, size_t ngates)
{
    const uint64_t amask = (uint64_t(1) << 63) - 1;
    for (size_t i = 0; i < ngates; i++)
    {
        const uint64_t a0 = gates[i].s0 & amask;
        const int most = gates[i].s1 == kNoSel ? 2 : 1;
        if (__builtin_popcountll(a0) > most || (n < 64 && (a0 >> n) != 0))
            return false;
        if (gates[i].s1 != kNoSel)
        {
            const uint64_t a1 = gates[i].s1 & amask;
            if (__builtin_popcountll(a1) > 1 || (n < 64 && (a1 >> n) != 0))
                return false;
        }
    }
    return true;
}

// theta * [S0][S1] with [x_j ^ beta] = beta + (1 - 2 beta) x_j, and for a
// one-selector gate on two bits theta * [x_u ^ x_v ^ beta]
// = theta * (beta + (1 - 2 beta) (x_u + x_v - 2 x_u x_v)).
void window_polynomial(unsigned n, const DiagGate *gates, size_t ngates, Poly &P)
{
    struct Lin
    {
        int bit;  // -1: constant selector
        double c; // constant part
        double s; // slope of x_bit
    };
    auto lin_of = [](uint64_t packed) -> Lin {
        const uint64_t alpha = packed & ((uint64_t(1) << 63) - 1);
        const bool beta = (packed >> 63) & 1u;
        if (alpha == 0)
            return {-1, beta ? 1.0 : 0.0, 0.0};
        return {__builtin_ctzll(alpha), beta ? 1.0 : 0.0, beta ? -1.0 : 1.0};
    };
    P.n = n;
    P.K = 0.0;
    P.L.assign(n, 0.0);
    P.Q.assign(size_t(n) * n, 0.0);
    for (size_t i = 0; i < ngates; i++)
    {
        const DiagGate &g = gates[i];
        const uint64_t a0 = g.s0 & ((uint64_t(1) << 63) - 1);
        if (g.s1 == kNoSel && __builtin_popcountll(a0) == 2)
        {
            const int u = __builtin_ctzll(a0), v = 63 - __builtin_clzll(a0);
            const double beta = (g.s0 >> 63) ? 1.0 : 0.0;
            const double sl = g.theta * (1.0 - 2.0 * beta);
            P.K += g.theta * beta;
            P.L[u] += sl;
            P.L[v] += sl;
            P.Q[size_t(u) * n + v] += -2.0 * sl;
            continue;
        }
        const Lin a = lin_of(g.s0);
        const Lin b = g.s1 == kNoSel ? Lin{-1, 1.0, 0.0} : lin_of(g.s1);
        const double th = g.theta;
        P.K += th * a.c * b.c;
        if (b.bit >= 0)
            P.L[b.bit] += th * a.c * b.s;
        if (a.bit >= 0)
            P.L[a.bit] += th * b.c * a.s;
        if (a.bit >= 0 && b.bit >= 0)
        {
            const double q = th * a.s * b.s;
            if (a.bit == b.bit)
                P.L[a.bit] += q;
            else
                P.Q[size_t(std::min(a.bit, b.bit)) * n + std::max(a.bit, b.bit)] += q;
        }
    }
}

int launch_block_coef(Ctx &ctx, const Poly &P, const int *inner, int nco, const int *outer, int nh, double *out)
{
    if (nco + 1 > kCoefStride)
    {
        set_error("block coefficients: too many inner bits");
        return EINVAL_;
    }
    std::vector<double> blk;
    blk.reserve(size_t(nco) * nh + nh + size_t(nh) * nh);
    for (int j = 0; j < nco; j++)
        for (int i = 0; i < nh; i++)
            blk.push_back(P.q(inner[j], outer[i]));
    for (int i = 0; i < nh; i++)
        blk.push_back(P.L[outer[i]]);
    for (int i = 0; i < nh; i++)
        for (int l = 0; l < nh; l++)
            blk.push_back(l > i ? P.q(outer[i], outer[l]) : 0.0);
    if (blk.empty())
        blk.push_back(0.0);
    void *d = nullptr;
    ZXA_TRY(ctx.ring.stage(blk.data(), blk.size() * sizeof(double), ctx.stream, &d));
    const double *dcq = static_cast<const double *>(d);
    const double *dhL = dcq + size_t(nco) * nh;
    const double *dhQ = dhL + nh;
    const uint64_t nblocks = uint64_t(1) << nh;
    const unsigned grid = static_cast<unsigned>((nblocks + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(diag_chunk_coef_kernel, dim3(grid), dim3(kThreads), 0, ctx.stream, dcq, dhL, dhQ, P.K, nco, nh,
                       nblocks, out);
    return hip_check(hipGetLastError(), "block coefficient kernel");
}

// The window's dispatch: the autotuned flat sweep, else the chunked
// polynomial sweep, else (windows that are not a degree-2 polynomial, or
// states narrower than a chunk) the per-gate general sweep.  ws (nullable)
// is a caller-given coefficient workspace of ws_bytes: a polynomial sweep
// whose per-chunk coefficients (128 bytes per chunk) do not fit it takes the
// polynomial sweep's larger chunks, and fails if those do not fit either.
// With ws null the context's scratch buffer grows as needed.
int diag_apply_ws(Ctx &ctx, void *state, int prec, unsigned n, const DiagGate *gates, size_t ngates, double *ws,
                  size_t ws_bytes)
{
    if (ngates == 0)
        return OK;
    if (n > 40)
    {
        set_error("diag_apply: state wider than 40 bits");
        return EINVAL_;
    }
    const bool simple = window_is_polynomial(n, gates, ngates);
    const auto fits = [&](unsigned chunk_bits) {
        return !ws || ((uint64_t(1) << (n - chunk_bits)) * kCoefStride * sizeof(double) <= ws_bytes);
    };
    const int ept = tune_get(prec == C64 ? "diag.ept.c64" : "diag.ept.c128");
    if (simple && ept > 0)
    {
        const unsigned need = (prec == C64 ? 7u : 6u) + 2u + (ept == 1 ? 0u : ept == 2 ? 1u : 2u);
        if (n >= need && (n - need) < 32 && fits(need))
            return diag_flat(ctx, state, prec, n, gates, ngates, ept,
                             tune_get(prec == C64 ? "diag.nt.c64" : "diag.nt.c128") != 0, ws);
    }
    const unsigned need = (prec == C64 ? Geo<true>::CB : Geo<false>::CB);
    if (simple && n >= need)
    {
        if (!fits(need))
        {
            set_error("diagonal sweep: the state's per-chunk coefficients (" +
                      std::to_string((uint64_t(1) << (n - need)) * kCoefStride * sizeof(double)) +
                      " bytes) exceed the workspace (" + std::to_string(ws_bytes) + " bytes)");
            return EINVAL_;
        }
        return diag_fast(ctx, state, prec, n, gates, ngates, ws);
    }
    // ring slot holds at most 256 KiB of gates; split very large windows
    const size_t per = ParamRing::kSlotBytes / sizeof(DiagGate);
    for (size_t off = 0; off < ngates; off += per)
        ZXA_TRY(diag_general(ctx, state, prec, n, gates + off, std::min(per, ngates - off)));
    return OK;
}

int diag_apply(Ctx &ctx, void *state, int prec, unsigned n, const DiagGate *gates, size_t ngates)
{
    return diag_apply_ws(ctx, state, prec, n, gates, ngates, nullptr, 0);
}

// ---- the workspace of the no-allocation entry point ----
//
// A file-scope device array: every GPU the library runs on holds one copy,
// at its own address, until the process ends.  Host code must not use the
// array's name as a device address; noalloc_workspace() looks the address up
// once per device (hipGetSymbolAddress, with that device current) and keeps it.
constexpr size_t kNoAllocBytes = size_t(32) << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

int noalloc_workspace(void **ptr, size_t *bytes)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    ZXA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
    {
        set_error("no-allocation workspace: device index out of range");
        return EINVAL_;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev])
    {
        void *p = nullptr;
        ZXA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
        {
            set_error("no-allocation workspace: symbol address is null");
            return EHIP;
        }
        addr[dev] = p;
    }
    *ptr = addr[dev];
    *bytes = kNoAllocBytes;
    return OK;
}

} // namespace zxa
