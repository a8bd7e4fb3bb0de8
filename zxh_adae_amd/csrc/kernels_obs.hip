// kernels_obs.hip -- diagonal Pauli-Z observables: expectation values,
// compensated norms, and the observable applied to a state.
//
// There is no counterpart in the reference (its engine measures by sampling
// only).  An observable H = sum_k c_k Z_{i_k} [Z_{j_k}] is diagonal in the
// computational basis; with s_q = 1 - 2 x_q its value on basis state x is a
// polynomial of degree <= 2 in the address bits,
//     h(x) = K + sum_b L_b x_b + sum_{b<c} Q_bc x_b x_c .
// The sweep splits the address like the diagonal phase sweep: lane bits
// (one 16-byte unit per thread, 256 threads), six row bits iterated by every
// thread in Gray-code order, and block bits, so that h costs one addition per
// amplitude.  Reductions accumulate in FP64 with Kahan compensation per
// thread, then a fixed-order tree per workgroup and a fixed-order final pass:
// the result is deterministic and, for complex64 states, carries FP64
// accuracy over the FP32 amplitudes.
#include "dev.h"
#include "poly.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace zxa
{
namespace
{

constexpr int kThreads = 256;
constexpr int kRowBits = 6;
constexpr int kRows = 1 << kRowBits;
constexpr int kCoefStride = 16; // matches launch_block_coef's layout

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

#define ZXA_CONST const __attribute__((address_space(4)))

template <bool kC64>
struct OGeo
{
    static constexpr int kEPU = kC64 ? 2 : 1;          // amplitudes per 16-byte unit
    static constexpr int kLane = 8 + (kC64 ? 1 : 0);  // lane element bits
    static constexpr int kBlock = kLane + kRowBits;    // block element bits
};

struct Kahan
{
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double v)
    {
        const double y = v - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Per-thread sums -> per-workgroup partial (fixed-order tree in LDS).
__device__ __forceinline__ void wg_reduce2(double a, double b, double *sh, double *out2)
{
    const int t = threadIdx.x;
    sh[t] = a;
    sh[kThreads + t] = b;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1)
    {
        if (t < w)
        {
            sh[t] += sh[t + w];
            sh[kThreads + t] += sh[kThreads + t + w];
        }
        __syncthreads();
    }
    if (t == 0)
    {
        out2[0] = sh[0];
        out2[1] = sh[kThreads];
    }
}

// ltab layout (host-built, per thread t and lane element e < kEPU):
//   A[t][e], d[t][e][j] (j < 6)      static lane part and lane-row couplings
//   then R[r] (r < 64)               static row part
// coef: per block, c_b for the kBlock inner bits and C at [kBlock]
// (launch_block_coef).  kApply: also write dst = h * src.
template <bool kC64, bool kApply>
__global__ __launch_bounds__(kThreads) void obs_kernel(const void *__restrict__ src, void *__restrict__ dst,
                                                       const double *__restrict__ ltab,
                                                       const double *__restrict__ coef, uint64_t nblocks,
                                                       double *__restrict__ partial)
{
    using G = OGeo<kC64>;
    constexpr int E = G::kEPU;
    __shared__ double sh[2 * kThreads];
    const int t = threadIdx.x;
    double A[E], d[E][kRowBits];
#pragma unroll
    for (int e = 0; e < E; e++)
    {
        const double *a = ltab + (size_t(t) * E + e) * (1 + kRowBits);
        A[e] = a[0];
#pragma unroll
        for (int j = 0; j < kRowBits; j++)
            d[e][j] = a[1 + j];
    }
    ZXA_CONST double *R = (ZXA_CONST double *)(ltab + size_t(kThreads) * E * (1 + kRowBits));
    Kahan eh, en;
    for (uint64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x)
    {
        ZXA_CONST double *c = (ZXA_CONST double *)(coef + blk * kCoefStride);
        double T[E], D[E][kRowBits];
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            // lane element bits: complex64 puts the in-unit bit first
            const int x = kC64 ? (t << 1) | e : t;
            double s = c[G::kBlock] + A[e];
#pragma unroll
            for (int b = 0; b < G::kLane; b++)
                if ((x >> b) & 1)
                    s += c[b];
            T[e] = s;
#pragma unroll
            for (int j = 0; j < kRowBits; j++)
                D[e][j] = d[e][j] + c[G::kLane + j];
        }
        const uint64_t base = blk << (G::kBlock - (kC64 ? 1 : 0)); // unit address of row 0
        double g[E];
#pragma unroll
        for (int e = 0; e < E; e++)
            g[e] = 0.0;
#pragma unroll 8
        for (int k = 0; k < kRows; k++)
        {
            const int r = k ^ (k >> 1);
            if (k > 0)
            {
                const int j = __builtin_ctz(k);
                const bool up = (r >> j) & 1;
#pragma unroll
                for (int e = 0; e < E; e++)
                    g[e] += up ? D[e][j] : -D[e][j];
            }
            const uint64_t u = base + (uint64_t(r) << 8) + t;
            if constexpr (kC64)
            {
                const fvec4 v = __builtin_nontemporal_load(static_cast<const fvec4 *>(src) + u);
                const double h0 = T[0] + R[r] + g[0];
                const double h1 = T[1] + R[r] + g[1];
                const double p0 = double(v.x) * double(v.x) + double(v.y) * double(v.y);
                const double p1 = double(v.z) * double(v.z) + double(v.w) * double(v.w);
                eh.add(p0 * h0);
                eh.add(p1 * h1);
                en.add(p0);
                en.add(p1);
                if constexpr (kApply)
                {
                    const float f0 = float(h0), f1 = float(h1);
                    __builtin_nontemporal_store(fvec4{v.x * f0, v.y * f0, v.z * f1, v.w * f1},
                                                static_cast<fvec4 *>(dst) + u);
                }
            }
            else
            {
                const dvec2 v = __builtin_nontemporal_load(static_cast<const dvec2 *>(src) + u);
                const double h = T[0] + R[r] + g[0];
                const double p = v.x * v.x + v.y * v.y;
                eh.add(p * h);
                en.add(p);
                if constexpr (kApply)
                    __builtin_nontemporal_store(dvec2{v.x * h, v.y * h}, static_cast<dvec2 *>(dst) + u);
            }
        }
    }
    wg_reduce2(eh.s, en.s, sh, partial + 2 * blockIdx.x);
}

// Batched: one block of one state per work item (states of different
// widths and observables in one launch); per-block partial sums.
struct OItem
{
    uint64_t first;   // first global block index of this state
    uint64_t src, dst; // dst = 0: expectation only
    uint64_t ltab;    // lane / row tables (as obs_kernel)
    uint64_t coef;    // block coefficients
};

template <bool kC64>
__global__ __launch_bounds__(kThreads) void obs_batch_kernel(const OItem *__restrict__ items, int nitems,
                                                             uint64_t total, double *__restrict__ partial)
{
    using G = OGeo<kC64>;
    constexpr int E = G::kEPU;
    __shared__ double sh[2 * kThreads];
    const int t = threadIdx.x;
    for (uint64_t g = blockIdx.x; g < total; g += gridDim.x)
    {
        ZXA_CONST OItem *IT = (ZXA_CONST OItem *)items;
        int lo = 0, hi = nitems - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (IT[mid].first <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        ZXA_CONST OItem &I = IT[lo];
        const uint64_t blk = g - I.first;
        const double *ltab = reinterpret_cast<const double *>(I.ltab);
        double A[E], d[E][kRowBits];
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            const double *a = ltab + (size_t(t) * E + e) * (1 + kRowBits);
            A[e] = a[0];
#pragma unroll
            for (int j = 0; j < kRowBits; j++)
                d[e][j] = a[1 + j];
        }
        ZXA_CONST double *R = (ZXA_CONST double *)(ltab + size_t(kThreads) * E * (1 + kRowBits));
        ZXA_CONST double *c = (ZXA_CONST double *)(reinterpret_cast<const double *>(I.coef) + blk * kCoefStride);
        double T[E], D[E][kRowBits];
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            const int x = kC64 ? (t << 1) | e : t;
            double s = c[G::kBlock] + A[e];
#pragma unroll
            for (int b = 0; b < G::kLane; b++)
                if ((x >> b) & 1)
                    s += c[b];
            T[e] = s;
#pragma unroll
            for (int j = 0; j < kRowBits; j++)
                D[e][j] = d[e][j] + c[G::kLane + j];
        }
        const uint64_t base = blk << (G::kBlock - (kC64 ? 1 : 0));
        const bool apply = I.dst != 0;
        Kahan eh, en;
        double gg[E];
#pragma unroll
        for (int e = 0; e < E; e++)
            gg[e] = 0.0;
#pragma unroll 8
        for (int k = 0; k < kRows; k++)
        {
            const int r = k ^ (k >> 1);
            if (k > 0)
            {
                const int j = __builtin_ctz(k);
                const bool up = (r >> j) & 1;
#pragma unroll
                for (int e = 0; e < E; e++)
                    gg[e] += up ? D[e][j] : -D[e][j];
            }
            const uint64_t u = base + (uint64_t(r) << 8) + t;
            if constexpr (kC64)
            {
                const fvec4 v = __builtin_nontemporal_load(reinterpret_cast<const fvec4 *>(I.src) + u);
                const double h0 = T[0] + R[r] + gg[0];
                const double h1 = T[1] + R[r] + gg[1];
                const double p0 = double(v.x) * double(v.x) + double(v.y) * double(v.y);
                const double p1 = double(v.z) * double(v.z) + double(v.w) * double(v.w);
                eh.add(p0 * h0);
                eh.add(p1 * h1);
                en.add(p0);
                en.add(p1);
                if (apply)
                {
                    const float f0 = float(h0), f1 = float(h1);
                    __builtin_nontemporal_store(fvec4{v.x * f0, v.y * f0, v.z * f1, v.w * f1},
                                                reinterpret_cast<fvec4 *>(I.dst) + u);
                }
            }
            else
            {
                const dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(I.src) + u);
                const double h = T[0] + R[r] + gg[0];
                const double p = v.x * v.x + v.y * v.y;
                eh.add(p * h);
                en.add(p);
                if (apply)
                    __builtin_nontemporal_store(dvec2{v.x * h, v.y * h}, reinterpret_cast<dvec2 *>(I.dst) + u);
            }
        }
        wg_reduce2(eh.s, en.s, sh, partial + 2 * g);
        __syncthreads();
    }
}

// Per state: its blocks' partials in order -> energies[2 i], energies[2 i + 1].
__global__ __launch_bounds__(64) void obs_batch_collect_kernel(const OItem *__restrict__ items, int nitems,
                                                               uint64_t total, const double *__restrict__ partial,
                                                               double *__restrict__ energies)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems)
        return;
    const uint64_t b0 = items[i].first;
    const uint64_t b1 = i + 1 < nitems ? items[i + 1].first : total;
    Kahan a, n;
    for (uint64_t b = b0; b < b1; b++)
    {
        a.add(partial[2 * b]);
        n.add(partial[2 * b + 1]);
    }
    energies[2 * i] = a.s;
    energies[2 * i + 1] = n.s;
}

// Small states (fewer address bits than one block): h evaluated per
// amplitude from the polynomial.  poly: K, L[n], Q[n*n].
template <bool kC64, bool kApply>
__global__ __launch_bounds__(kThreads) void obs_small_kernel(const void *__restrict__ src, void *__restrict__ dst,
                                                             const double *__restrict__ poly, int n,
                                                             double *__restrict__ partial)
{
    __shared__ double sh[2 * kThreads];
    Kahan eh, en;
    const uint64_t N = uint64_t(1) << n;
    for (uint64_t x = blockIdx.x * uint64_t(kThreads) + threadIdx.x; x < N; x += uint64_t(gridDim.x) * kThreads)
    {
        double h = poly[0];
        for (int b = 0; b < n; b++)
        {
            if (!((x >> b) & 1))
                continue;
            h += poly[1 + b];
            for (int c2 = b + 1; c2 < n; c2++)
                if ((x >> c2) & 1)
                    h += poly[1 + n + b * n + c2];
        }
        double p;
        if constexpr (kC64)
        {
            const float2 v = static_cast<const float2 *>(src)[x];
            p = double(v.x) * double(v.x) + double(v.y) * double(v.y);
            if constexpr (kApply)
                static_cast<float2 *>(dst)[x] = make_float2(v.x * float(h), v.y * float(h));
        }
        else
        {
            const double2 v = static_cast<const double2 *>(src)[x];
            p = v.x * v.x + v.y * v.y;
            if constexpr (kApply)
                static_cast<double2 *>(dst)[x] = make_double2(v.x * h, v.y * h);
        }
        eh.add(p * h);
        en.add(p);
    }
    wg_reduce2(eh.s, en.s, sh, partial + 2 * blockIdx.x);
}

// Fixed-order final reduction of the per-workgroup partials (one workgroup).
__global__ __launch_bounds__(kThreads) void partial_sum_kernel(const double *__restrict__ partial, int count,
                                                               double *__restrict__ out2)
{
    __shared__ double sh[2 * kThreads];
    Kahan a, b;
    for (int i = threadIdx.x; i < count; i += kThreads)
    {
        a.add(partial[2 * i]);
        b.add(partial[2 * i + 1]);
    }
    wg_reduce2(a.s, b.s, sh, out2);
}

// Observable polynomial from Z / ZZ terms (j < 0: single Z).
void observable_poly(unsigned n, const int32_t *qi, const int32_t *qj, const double *coef, size_t nterms,
                     Poly &P)
{
    P.n = n;
    P.K = 0.0;
    P.L.assign(n, 0.0);
    P.Q.assign(size_t(n) * n, 0.0);
    for (size_t k = 0; k < nterms; k++)
    {
        const double c = coef[k];
        const int i = qi[k], j = qj[k];
        if (j < 0 || j == i)
        {
            // Z_i = 1 - 2 x_i  (Z_i Z_i = 1)
            if (j == i)
            {
                P.K += c;
                continue;
            }
            P.K += c;
            P.L[i] -= 2.0 * c;
            continue;
        }
        // Z_i Z_j = 1 - 2 x_i - 2 x_j + 4 x_i x_j
        P.K += c;
        P.L[i] -= 2.0 * c;
        P.L[j] -= 2.0 * c;
        const int a = std::min(i, j), b = std::max(i, j);
        P.Q[size_t(a) * n + b] += 4.0 * c;
    }
}

// Two states of 2^n amplitudes compared in one read of each (a: kA64 ? c64 :
// c128, b likewise): per thread, Kahan sums in FP64 of Re and Im <a|b> =
// sum conj(a_i) b_i, |a|^2 and |b|^2, and the running max of |a_i - b_i|^2
// (order-free); then the fixed-order workgroup tree of wg_reduce2 on pairs.
// Each thread takes one pair of amplitudes per step (a 16-byte load of a c64
// state, two of a c128 state).  partial: 6 doubles per workgroup.
template <bool kA64, bool kB64>
__global__ __launch_bounds__(kThreads) void compare_kernel(const void *__restrict__ a, const void *__restrict__ b,
                                                           uint64_t npairs, double *__restrict__ partial)
{
    __shared__ double sh[2 * kThreads];
    Kahan re, im, na, nb;
    double mx = 0.0;
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < npairs; i += stride)
    {
        double av[4], bv[4];
        if (kA64)
        {
            const fvec4 v = __builtin_nontemporal_load(static_cast<const fvec4 *>(a) + i);
            av[0] = v.x, av[1] = v.y, av[2] = v.z, av[3] = v.w;
        }
        else
        {
            const dvec2 v0 = __builtin_nontemporal_load(static_cast<const dvec2 *>(a) + 2 * i);
            const dvec2 v1 = __builtin_nontemporal_load(static_cast<const dvec2 *>(a) + 2 * i + 1);
            av[0] = v0.x, av[1] = v0.y, av[2] = v1.x, av[3] = v1.y;
        }
        if (kB64)
        {
            const fvec4 v = __builtin_nontemporal_load(static_cast<const fvec4 *>(b) + i);
            bv[0] = v.x, bv[1] = v.y, bv[2] = v.z, bv[3] = v.w;
        }
        else
        {
            const dvec2 v0 = __builtin_nontemporal_load(static_cast<const dvec2 *>(b) + 2 * i);
            const dvec2 v1 = __builtin_nontemporal_load(static_cast<const dvec2 *>(b) + 2 * i + 1);
            bv[0] = v0.x, bv[1] = v0.y, bv[2] = v1.x, bv[3] = v1.y;
        }
#pragma unroll
        for (int e = 0; e < 2; e++)
        {
            const double ar = av[2 * e], ai = av[2 * e + 1], br = bv[2 * e], bi = bv[2 * e + 1];
            re.add(ar * br + ai * bi);
            im.add(ar * bi - ai * br);
            na.add(ar * ar + ai * ai);
            nb.add(br * br + bi * bi);
            const double dr = ar - br, di = ai - bi;
            mx = fmax(mx, dr * dr + di * di);
        }
    }
    double *out = partial + 6 * size_t(blockIdx.x);
    double two[2];
    wg_reduce2(re.s, im.s, sh, two);
    if (threadIdx.x == 0)
        out[0] = two[0], out[1] = two[1];
    __syncthreads();
    wg_reduce2(na.s, nb.s, sh, two);
    if (threadIdx.x == 0)
        out[2] = two[0], out[3] = two[1];
    __syncthreads();
    // max: tree of fmax (exact in any order)
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w)
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[4] = sh[0], out[5] = 0.0;
}

// Fixed-order final pass over compare_kernel's partials (one workgroup).
__global__ __launch_bounds__(kThreads) void compare_final_kernel(const double *__restrict__ partial, int count,
                                                                 double *__restrict__ out)
{
    __shared__ double sh[2 * kThreads];
    Kahan re, im, na, nb;
    double mx = 0.0;
    for (int i = threadIdx.x; i < count; i += kThreads)
    {
        const double *p = partial + 6 * size_t(i);
        re.add(p[0]);
        im.add(p[1]);
        na.add(p[2]);
        nb.add(p[3]);
        mx = fmax(mx, p[4]);
    }
    double two[2];
    wg_reduce2(re.s, im.s, sh, two);
    if (threadIdx.x == 0)
        out[0] = two[0], out[1] = two[1];
    __syncthreads();
    wg_reduce2(na.s, nb.s, sh, two);
    if (threadIdx.x == 0)
        out[2] = two[0], out[3] = two[1];
    __syncthreads();
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w)
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[4] = sqrt(sh[0]);
}

template <typename K>
int grid_of(Ctx &ctx, K kernel, uint64_t items)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const uint64_t cap = uint64_t(per_cu) * uint64_t(ctx.num_cus > 0 ? ctx.num_cus : 256);
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(items, cap)));
}

} // namespace

int observable_apply(Ctx &ctx, const void *src, void *dst, int prec, unsigned n, const int32_t *qi,
                     const int32_t *qj, const double *coef, size_t nterms, double *out2)
{
    const bool c64 = prec == C64;
    const bool apply = dst != nullptr;
    for (size_t k = 0; k < nterms; k++)
        if (qi[k] < 0 || unsigned(qi[k]) >= n || qj[k] >= int(n))
        {
            set_error("observable: qubit index outside the state");
            return EINVAL_;
        }
    if (n > 40)
    {
        set_error("observable: state wider than 40 bits");
        return EINVAL_;
    }
    Poly P;
    observable_poly(n, qi, qj, coef, nterms, P);
    const unsigned bb = c64 ? OGeo<true>::kBlock : OGeo<false>::kBlock;
    const uint64_t bytes = (apply ? 2ull : 1ull) * (uint64_t(1) << n) * elem_bytes(prec);
    int grid = 0;
    if (n < bb)
    {
        std::vector<double> pt(1 + n + size_t(n) * n, 0.0);
        pt[0] = P.K;
        for (unsigned b = 0; b < n; b++)
        {
            pt[1 + b] = P.L[b];
            for (unsigned c = b + 1; c < n; c++)
                pt[1 + n + b * n + c] = P.q(int(b), int(c));
        }
        void *dpt = nullptr;
        ZXA_TRY(ctx.ring.stage(pt.data(), pt.size() * sizeof(double), ctx.stream, &dpt));
        const uint64_t N = uint64_t(1) << n;
        grid = int(std::max<uint64_t>(1, std::min<uint64_t>((N + kThreads - 1) / kThreads, 1024)));
        ZXA_TRY(ctx.scratch_e.ensure(size_t(grid) * 2 * sizeof(double) + 2 * sizeof(double)));
        double *part = static_cast<double *>(ctx.scratch_e.ptr);
        ZXA_TRY(ctx.tic());
        auto k = c64 ? (apply ? obs_small_kernel<true, true> : obs_small_kernel<true, false>)
                     : (apply ? obs_small_kernel<false, true> : obs_small_kernel<false, false>);
        hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, src, dst, static_cast<const double *>(dpt),
                           int(n), part);
        ZXA_HIP(hipGetLastError());
    }
    else
    {
        // lane table (per thread and lane element) and row table
        const int E = c64 ? 2 : 1;
        const int LB = c64 ? OGeo<true>::kLane : OGeo<false>::kLane;
        std::vector<double> lt(size_t(kThreads) * E * (1 + kRowBits) + kRows, 0.0);
        for (int t = 0; t < kThreads; t++)
            for (int e = 0; e < E; e++)
            {
                const int x = c64 ? (t << 1) | e : t;
                double *a = lt.data() + (size_t(t) * E + e) * (1 + kRowBits);
                double s = 0.0;
                for (int b = 0; b < LB; b++)
                {
                    if (!((x >> b) & 1))
                        continue;
                    s += P.L[b];
                    for (int c = b + 1; c < LB; c++)
                        if ((x >> c) & 1)
                            s += P.q(b, c);
                }
                a[0] = s;
                for (int j = 0; j < kRowBits; j++)
                {
                    double dj = 0.0;
                    for (int b = 0; b < LB; b++)
                        if ((x >> b) & 1)
                            dj += P.q(b, LB + j);
                    a[1 + j] = dj;
                }
            }
        double *R = lt.data() + size_t(kThreads) * E * (1 + kRowBits);
        for (int r = 0; r < kRows; r++)
        {
            double s = 0.0;
            for (int j = 0; j < kRowBits; j++)
            {
                if (!((r >> j) & 1))
                    continue;
                s += P.L[LB + j];
                for (int k = j + 1; k < kRowBits; k++)
                    if ((r >> k) & 1)
                        s += P.q(LB + j, LB + k);
            }
            R[r] = s;
        }
        void *dlt = nullptr;
        ZXA_TRY(ctx.ring.stage(lt.data(), lt.size() * sizeof(double), ctx.stream, &dlt));
        const int nco = int(bb);
        const int nh = int(n) - nco;
        std::vector<int> inner(nco), outer(nh);
        for (int i = 0; i < nco; i++)
            inner[i] = i;
        for (int i = 0; i < nh; i++)
            outer[i] = nco + i;
        const uint64_t nblocks = uint64_t(1) << nh;
        ZXA_TRY(ctx.scratch_a.ensure(nblocks * kCoefStride * sizeof(double)));
        double *coefs = static_cast<double *>(ctx.scratch_a.ptr);
        ZXA_TRY(launch_block_coef(ctx, P, inner.data(), nco, outer.data(), nh, coefs));
        auto k = c64 ? (apply ? obs_kernel<true, true> : obs_kernel<true, false>)
                     : (apply ? obs_kernel<false, true> : obs_kernel<false, false>);
        grid = grid_of(ctx, k, nblocks);
        ZXA_TRY(ctx.scratch_e.ensure(size_t(grid) * 2 * sizeof(double) + 2 * sizeof(double)));
        double *part = static_cast<double *>(ctx.scratch_e.ptr);
        ZXA_TRY(ctx.tic());
        hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, src, dst, static_cast<const double *>(dlt),
                           static_cast<const double *>(coefs), nblocks, part);
        ZXA_HIP(hipGetLastError());
    }
    double *part = static_cast<double *>(ctx.scratch_e.ptr);
    double *res = part + size_t(grid) * 2;
    hipLaunchKernelGGL(partial_sum_kernel, dim3(1), dim3(kThreads), 0, ctx.stream, part, grid, res);
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_OBS, bytes));
    if (out2)
    {
        ZXA_HIP(hipMemcpyAsync(out2, res, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
        ZXA_HIP(hipStreamSynchronize(ctx.stream));
    }
    return ctx.ring.done(ctx.stream);
}

int compare_states(Ctx &ctx, const void *a, int prec_a, const void *b, int prec_b, unsigned n, double *out5)
{
    if (n < 1 || n > 40)
    {
        set_error("compare: state width must be 1..40 bits");
        return EINVAL_;
    }
    const uint64_t npairs = uint64_t(1) << (n - 1);
    const bool a64 = prec_a == C64, b64 = prec_b == C64;
    auto k = a64 ? (b64 ? compare_kernel<true, true> : compare_kernel<true, false>)
                 : (b64 ? compare_kernel<false, true> : compare_kernel<false, false>);
    const int grid = grid_of(ctx, k, (npairs + kThreads - 1) / kThreads);
    ZXA_TRY(ctx.scratch_e.ensure(size_t(grid) * 6 * sizeof(double) + 8 * sizeof(double)));
    double *part = static_cast<double *>(ctx.scratch_e.ptr);
    double *res = part + size_t(grid) * 6;
    const uint64_t bytes = (uint64_t(1) << n) * (elem_bytes(prec_a) + elem_bytes(prec_b));
    ZXA_TRY(ctx.tic());
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, a, b, npairs, part);
    ZXA_HIP(hipGetLastError());
    hipLaunchKernelGGL(compare_final_kernel, dim3(1), dim3(kThreads), 0, ctx.stream, part, grid, res);
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_OBS, bytes));
    ZXA_HIP(hipMemcpyAsync(out5, res, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    ZXA_HIP(hipStreamSynchronize(ctx.stream));
    return OK;
}

ObsBatch::~ObsBatch()
{
    for (void *p : bufs)
        if (p)
            (void)hipFree(p);
}

namespace
{
// Lane and row tables of obs_kernel for one observable polynomial.
std::vector<double> lane_row_tables(bool c64, const Poly &P)
{
    const int E = c64 ? 2 : 1;
    const int LB = c64 ? OGeo<true>::kLane : OGeo<false>::kLane;
    std::vector<double> lt(size_t(kThreads) * E * (1 + kRowBits) + kRows, 0.0);
    for (int t = 0; t < kThreads; t++)
        for (int e = 0; e < E; e++)
        {
            const int x = c64 ? (t << 1) | e : t;
            double *a = lt.data() + (size_t(t) * E + e) * (1 + kRowBits);
            double s = 0.0;
            for (int b = 0; b < LB; b++)
            {
                if (!((x >> b) & 1))
                    continue;
                s += P.L[b];
                for (int c = b + 1; c < LB; c++)
                    if ((x >> c) & 1)
                        s += P.q(b, c);
            }
            a[0] = s;
            for (int j = 0; j < kRowBits; j++)
            {
                double dj = 0.0;
                for (int b = 0; b < LB; b++)
                    if ((x >> b) & 1)
                        dj += P.q(b, LB + j);
                a[1 + j] = dj;
            }
        }
    double *R = lt.data() + size_t(kThreads) * E * (1 + kRowBits);
    for (int r = 0; r < kRows; r++)
    {
        double s = 0.0;
        for (int j = 0; j < kRowBits; j++)
        {
            if (!((r >> j) & 1))
                continue;
            s += P.L[LB + j];
            for (int k = j + 1; k < kRowBits; k++)
                if ((r >> k) & 1)
                    s += P.q(LB + j, LB + k);
        }
        R[r] = s;
    }
    return lt;
}
} // namespace

int obs_batch_prepare(Ctx &ctx, int prec, const std::vector<ObsSpec> &specs, double *d_energies, ObsBatch &b)
{
    const bool c64 = prec == C64;
    b.prec = prec;
    b.nitems = specs.size();
    const unsigned bb = c64 ? OGeo<true>::kBlock : OGeo<false>::kBlock;
    std::vector<OItem> items(specs.size());
    uint64_t first = 0;
    for (size_t i = 0; i < specs.size(); i++)
    {
        const ObsSpec &sp = specs[i];
        if (sp.n < bb || sp.n > 40)
        {
            set_error("observable batch: state narrower than one block (14 bits complex128, 15 complex64)");
            return EINVAL_;
        }
        Poly P;
        observable_poly(sp.n, sp.qi.data(), sp.qj.data(), sp.coef.data(), sp.coef.size(), P);
        const std::vector<double> lt = lane_row_tables(c64, P);
        void *dlt = nullptr;
        ZXA_HIP(hipMalloc(&dlt, lt.size() * sizeof(double)));
        b.bufs.push_back(dlt);
        ZXA_HIP(hipMemcpy(dlt, lt.data(), lt.size() * sizeof(double), hipMemcpyHostToDevice));
        const int nco = int(bb), nh = int(sp.n) - nco;
        std::vector<int> inner(nco), outer(nh);
        for (int k = 0; k < nco; k++)
            inner[k] = k;
        for (int k = 0; k < nh; k++)
            outer[k] = nco + k;
        const uint64_t nblocks = uint64_t(1) << nh;
        void *dco = nullptr;
        ZXA_HIP(hipMalloc(&dco, nblocks * kCoefStride * sizeof(double)));
        b.bufs.push_back(dco);
        ZXA_TRY(launch_block_coef(ctx, P, inner.data(), nco, outer.data(), nh, static_cast<double *>(dco)));
        ZXA_TRY(ctx.ring.done(ctx.stream));
        items[i].first = first;
        items[i].src = reinterpret_cast<uint64_t>(sp.src);
        items[i].dst = reinterpret_cast<uint64_t>(sp.dst);
        items[i].ltab = reinterpret_cast<uint64_t>(dlt);
        items[i].coef = reinterpret_cast<uint64_t>(dco);
        first += nblocks;
    }
    b.total_blocks = first;
    ZXA_HIP(hipMalloc(&b.d_items, std::max<size_t>(1, items.size()) * sizeof(OItem)));
    b.bufs.push_back(b.d_items);
    ZXA_HIP(hipMemcpy(b.d_items, items.data(), items.size() * sizeof(OItem), hipMemcpyHostToDevice));
    ZXA_HIP(hipMalloc(&b.d_partial, std::max<uint64_t>(1, first) * 2 * sizeof(double)));
    b.bufs.push_back(b.d_partial);
    b.d_tables = d_energies;
    b.grid = grid_of(ctx, c64 ? obs_batch_kernel<true> : obs_batch_kernel<false>, first);
    return hip_check(hipStreamSynchronize(ctx.stream), "observable batch prepare");
}

int obs_batch_launch(Ctx &ctx, ObsBatch &b)
{
    if (!b.nitems)
        return OK;
    auto k = b.prec == C64 ? obs_batch_kernel<true> : obs_batch_kernel<false>;
    hipLaunchKernelGGL(k, dim3(b.grid), dim3(kThreads), 0, ctx.stream, static_cast<const OItem *>(b.d_items),
                       int(b.nitems), b.total_blocks, static_cast<double *>(b.d_partial));
    ZXA_HIP(hipGetLastError());
    const unsigned g2 = unsigned((b.nitems + 63) / 64);
    hipLaunchKernelGGL(obs_batch_collect_kernel, dim3(g2), dim3(64), 0, ctx.stream,
                       static_cast<const OItem *>(b.d_items), int(b.nitems), b.total_blocks,
                       static_cast<const double *>(b.d_partial), static_cast<double *>(b.d_tables));
    return hip_check(hipGetLastError(), "observable batch launch");
}

} // namespace zxa
