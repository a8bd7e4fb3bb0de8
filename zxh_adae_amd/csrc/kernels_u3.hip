// kernels_u3.hip -- single-qubit batches, permutations and CP fallback.
//
// The reference applies a fused U3 cluster as one full-state sweep per
// descriptor (u3_block_batch_kernel_cuda, reference
// src/zxh/src/kernel/kernel_cuda.cu:390-423; CPU kernel_general.cpp:173-202).
// Here a cluster is split into tile passes.  A tile is 2^12 amplitudes whose
// addresses differ in 12 chosen bits: the four lowest address bits (so that
// every 16-lane group reads 256 contiguous bytes) plus up to eight target
// qubits.  Each workgroup loads a tile, applies every gate whose target lies
// in the tile, and writes it back, so one HBM sweep applies up to 12 gates.
//
// Inside a tile each thread holds 16 amplitudes.  Gates act on register
// pairs; to reach the other tile bits the tile is transposed through LDS (an
// XOR-swizzled buffer, conflict-free for every access pattern used) into a
// layout whose register bits are those tile bits.  Passes whose targets all
// sit in the register bits of the load layout touch no LDS.
//
// Gates whose pair offset delta has several bits (CX transport in the ZXH
// affine mode), and states narrower than a tile, use the pair kernel: one
// sweep per gate, pairs enumerated by inserting a zero at the lowest delta bit.
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
constexpr int kTileBits = 12;
constexpr int kRegs = 16;

struct cd
{
    double x, y;
};

__device__ __forceinline__ bool sel_eval(uint64_t packed, uint64_t x)
{
    const uint64_t alpha = packed & ((uint64_t(1) << 63) - 1);
    return ((__popcll(x & alpha) & 1u) ^ static_cast<unsigned>(packed >> 63)) != 0u;
}

// Insert a zero bit at position p.
__host__ __device__ __forceinline__ uint64_t insert0(uint64_t x, unsigned p)
{
    const uint64_t lo = x & ((uint64_t(1) << p) - 1);
    return ((x >> p) << (p + 1)) | lo;
}

struct TileGate
{
    double u[8];         // m00 re,im  m01  m10  m11 (already swapped when beta flips it)
    uint64_t alpha_rest; // selector bits other than the target (per-pair flip)
    uint32_t beta;
    uint32_t present;
};

struct TilePass
{
    uint32_t tb[kTileBits]; // global bit positions of the tile bits, ascending
    uint32_t pad;
    uint64_t ntiles;
    TileGate g[kTileBits]; // by tile bit index
};

// Register bits of the three layouts (tile bit indices).
template <bool kC64>
struct Lay;
template <>
struct Lay<false>
{
    static constexpr int R[3][4] = {{8, 9, 10, 11}, {4, 5, 6, 7}, {0, 1, 2, 3}};
};
template <>
struct Lay<true>
{
    static constexpr int R[3][4] = {{0, 9, 10, 11}, {1, 2, 3, 4}, {5, 6, 7, 8}};
};

template <bool kC64, int L>
__device__ constexpr int thread_tile_bit(int j)
{
    // j-th (ascending) tile bit that is not a register bit of layout L
    int cnt = 0;
    for (int b = 0; b < kTileBits; b++)
    {
        bool reg = false;
        for (int i = 0; i < 4; i++)
            reg = reg || (Lay<kC64>::R[L][i] == b);
        if (reg)
            continue;
        if (cnt == j)
            return b;
        cnt++;
    }
    return -1;
}

template <bool kC64, int L>
__device__ constexpr int reg_tau(int r)
{
    int t = 0;
    for (int i = 0; i < 4; i++)
        if ((r >> i) & 1)
            t |= 1 << Lay<kC64>::R[L][i];
    return t;
}

template <bool kC64, int L>
__device__ __forceinline__ int thread_tau(int t)
{
    int tau = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
        if ((t >> j) & 1)
            tau |= 1 << thread_tile_bit<kC64, L>(j);
    return tau;
}

template <bool kC64, int L>
__device__ __forceinline__ uint64_t thread_goff(int t, const uint32_t *tb)
{
    uint64_t off = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
        if ((t >> j) & 1)
            off |= uint64_t(1) << tb[thread_tile_bit<kC64, L>(j)];
    return off;
}

template <bool kC64, int L>
__device__ __forceinline__ uint64_t reg_goff(int r, const uint32_t *tb)
{
    uint64_t off = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
        if ((r >> i) & 1)
            off |= uint64_t(1) << tb[Lay<kC64>::R[L][i]];
    return off;
}

// LDS index swizzle: conflict-free for all three layouts.
template <bool kC64>
__device__ __forceinline__ int swz(int tau)
{
    if constexpr (kC64)
        return tau ^ (((tau >> 5) & 15) << 1);
    else
        return tau ^ ((tau >> 4) & 15);
}

template <bool kC64>
struct Elem;
template <>
struct Elem<false>
{
    using T = double2;
    using R = double;
};
template <>
struct Elem<true>
{
    using T = float2;
    using R = float;
};

// Apply the gates of layout L to the 16 registers.
template <bool kC64, int L>
__device__ __forceinline__ void apply_layout(typename Elem<kC64>::T (&v)[kRegs], const TilePass &P, uint64_t gbase,
                                             uint64_t gthr)
{
    using R = typename Elem<kC64>::R;
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const TileGate &g = P.g[Lay<kC64>::R[L][i]];
        if (!g.present)
            continue;
        const R m00x = R(g.u[0]), m00y = R(g.u[1]), m01x = R(g.u[2]), m01y = R(g.u[3]);
        const R m10x = R(g.u[4]), m10y = R(g.u[5]), m11x = R(g.u[6]), m11y = R(g.u[7]);
        const bool per_pair = g.alpha_rest != 0;
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            if ((r >> i) & 1)
                continue;
            const int r1 = r | (1 << i);
            const typename Elem<kC64>::T a = v[r], b = v[r1];
            bool flip = false;
            if (per_pair)
            {
                const uint64_t gi = gbase + gthr + reg_goff<kC64, L>(r, P.tb);
                flip = ((__popcll(gi & g.alpha_rest) & 1u) ^ g.beta) != 0u;
            }
            // (a, b) -> (m00 a + m01 b, m10 a + m11 b); flipped: even side is b
            const R p00x = flip ? m11x : m00x, p00y = flip ? m11y : m00y;
            const R p01x = flip ? m10x : m01x, p01y = flip ? m10y : m01y;
            const R p10x = flip ? m01x : m10x, p10y = flip ? m01y : m10y;
            const R p11x = flip ? m00x : m11x, p11y = flip ? m00y : m11y;
            typename Elem<kC64>::T na, nb;
            na.x = p00x * a.x - p00y * a.y + p01x * b.x - p01y * b.y;
            na.y = p00x * a.y + p00y * a.x + p01x * b.y + p01y * b.x;
            nb.x = p10x * a.x - p10y * a.y + p11x * b.x - p11y * b.y;
            nb.y = p10x * a.y + p10y * a.x + p11x * b.y + p11y * b.x;
            v[r] = na;
            v[r1] = nb;
        }
    }
}

template <bool kC64, int LW, int LR>
__device__ __forceinline__ void transpose(typename Elem<kC64>::T (&v)[kRegs], typename Elem<kC64>::T *lds, int t)
{
    const int tw = thread_tau<kC64, LW>(t);
    const int tr = thread_tau<kC64, LR>(t);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        lds[swz<kC64>(tw | reg_tau<kC64, LW>(r))] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        v[r] = lds[swz<kC64>(tr | reg_tau<kC64, LR>(r))];
}

__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cd cis(double a)
{
    double s, c;
    sincos(a, &s, &c);
    return {c, s};
}
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffu), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                static_cast<unsigned>(lo));
}

// Diagonal epilogue of a fused pass, applied in the load/store layout after
// every gate of the pass.  With theta the window's phase polynomial and the
// tile's element index split into thread bits T and register bits R,
//     theta = A(t) + B(r) + sum_{j in R} r_j d_j(t)
// where A is per thread and tile, B uniform per tile (computed by lane r),
// and d_j tile-independent.  The 16 registers are visited in Gray-code order
// so that exp(i sum r_j d_j) costs one complex product per element.
// dtab: Lt[12] | Qt[12 * 12] (linear and pairwise coefficients of the tile
// bits); coef: per tile, c_i (i < 12) and C from the non-tile bits.
template <bool kC64>
struct DiagEpi
{
    double thr_const = 0.0; // sum over set thread bits of Lt and pairwise Qt
    double reg_const = 0.0; // same over the register bits of lane r = lane & 15
    cd g[4];                // exp(i d_j(t)) for the four register bits

    __device__ __forceinline__ void init(const double *__restrict__ dtab, int t)
    {
        const double *Lt = dtab;
        const double *Qt = dtab + kTileBits;
#pragma unroll
        for (int j = 0; j < 8; j++)
        {
            if (!((t >> j) & 1))
                continue;
            const int i = thread_tile_bit<kC64, 0>(j);
            thr_const += Lt[i];
#pragma unroll
            for (int k = j + 1; k < 8; k++)
                if ((t >> k) & 1)
                    thr_const += Qt[i * kTileBits + thread_tile_bit<kC64, 0>(k)];
        }
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
        {
            const int R = Lay<kC64>::R[0][jj];
            double d = 0.0;
#pragma unroll
            for (int j = 0; j < 8; j++)
            {
                if (!((t >> j) & 1))
                    continue;
                const int i = thread_tile_bit<kC64, 0>(j);
                d += Qt[(i < R ? i : R) * kTileBits + (i < R ? R : i)];
            }
            g[jj] = cis(d);
        }
        const int r = t & 15;
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
        {
            if (!((r >> jj) & 1))
                continue;
            reg_const += Lt[Lay<kC64>::R[0][jj]];
#pragma unroll
            for (int kk = jj + 1; kk < 4; kk++)
                if ((r >> kk) & 1)
                    reg_const += Qt[Lay<kC64>::R[0][jj] * kTileBits + Lay<kC64>::R[0][kk]];
        }
    }

    __device__ __forceinline__ void apply(typename Elem<kC64>::T (&v)[kRegs], const double *__restrict__ cc, int t)
    {
        double A = thr_const + cc[kTileBits];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if ((t >> j) & 1)
                A += cc[thread_tile_bit<kC64, 0>(j)];
        double B = reg_const;
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
            if (((t & 15) >> jj) & 1)
                B += cc[Lay<kC64>::R[0][jj]];
        const cd bph = cis(B);
        cd w = cis(A);
#pragma unroll
        for (int k = 0; k < kRegs; k++)
        {
            const int r = k ^ (k >> 1);
            if (k > 0)
            {
                const int j = __builtin_ctz(k);
                const cd gj = ((r >> j) & 1) ? g[j] : cd{g[j].x, -g[j].y};
                w = cmul(w, gj);
            }
            const cd ph = cmul(w, {readlane_d(bph.x, r), readlane_d(bph.y, r)});
            if constexpr (kC64)
            {
                const float px = static_cast<float>(ph.x), py = static_cast<float>(ph.y);
                const float2 a = v[r];
                v[r] = make_float2(a.x * px - a.y * py, a.x * py + a.y * px);
            }
            else
            {
                const double2 a = v[r];
                v[r] = make_double2(a.x * ph.x - a.y * ph.y, a.x * ph.y + a.y * ph.x);
            }
        }
    }
};

template <bool kC64, bool kL1, bool kL2, bool kDiag>
__global__ __launch_bounds__(kThreads) void u3_tile_kernel(void *__restrict__ state, const TilePass P,
                                                           const double *__restrict__ dtab,
                                                           const double *__restrict__ dcoef)
{
    using T = typename Elem<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? (1 << kTileBits) : 1];
    const int t = threadIdx.x;
    const uint64_t g0 = thread_goff<kC64, 0>(t, P.tb);
    const uint64_t g1 = thread_goff<kC64, 1>(t, P.tb);
    const uint64_t g2 = thread_goff<kC64, 2>(t, P.tb);
    T *p = static_cast<T *>(state);
    DiagEpi<kC64> epi;
    if constexpr (kDiag)
        epi.init(dtab, t);

    for (uint64_t tile = blockIdx.x; tile < P.ntiles; tile += gridDim.x)
    {
        uint64_t base = tile;
#pragma unroll
        for (int k = 0; k < kTileBits; k++)
            base = insert0(base, P.tb[k]);

        T v[kRegs];
        if constexpr (!kC64)
        {
#pragma unroll
            for (int r = 0; r < kRegs; r++)
                v[r] = p[base + g0 + reg_goff<kC64, 0>(r, P.tb)];
        }
        else
        {
            // register bit 0 is tile bit 0 (= address bit 0): 16-byte pairs
            const float4 *p4 = static_cast<const float4 *>(state);
#pragma unroll
            for (int r = 0; r < kRegs; r += 2)
            {
                const float4 w = p4[(base + g0 + reg_goff<kC64, 0>(r, P.tb)) >> 1];
                v[r] = make_float2(w.x, w.y);
                v[r + 1] = make_float2(w.z, w.w);
            }
        }

        apply_layout<kC64, 0>(v, P, base, g0);
        if constexpr (kL1)
        {
            transpose<kC64, 0, 1>(v, lds, t);
            apply_layout<kC64, 1>(v, P, base, g1);
            if constexpr (kL2)
            {
                transpose<kC64, 1, 2>(v, lds, t);
                apply_layout<kC64, 2>(v, P, base, g2);
                transpose<kC64, 2, 0>(v, lds, t);
            }
            else
            {
                transpose<kC64, 1, 0>(v, lds, t);
            }
        }
        else if constexpr (kL2)
        {
            transpose<kC64, 0, 2>(v, lds, t);
            apply_layout<kC64, 2>(v, P, base, g2);
            transpose<kC64, 2, 0>(v, lds, t);
        }
        if constexpr (kDiag)
            epi.apply(v, dcoef + tile * 16, t);

        if constexpr (!kC64)
        {
#pragma unroll
            for (int r = 0; r < kRegs; r++)
                p[base + g0 + reg_goff<kC64, 0>(r, P.tb)] = v[r];
        }
        else
        {
            float4 *p4 = static_cast<float4 *>(state);
#pragma unroll
            for (int r = 0; r < kRegs; r += 2)
                p4[(base + g0 + reg_goff<kC64, 0>(r, P.tb)) >> 1] = make_float4(v[r].x, v[r].y, v[r + 1].x, v[r + 1].y);
        }
    }
}

// General pair sweep: every i with S(i) = 0 is paired with i ^ delta.
template <bool kC64>
__global__ __launch_bounds__(kThreads) void u3_pair_kernel(void *__restrict__ state, uint64_t npairs, uint64_t delta,
                                                           unsigned pivot, uint64_t packed, double u00x, double u00y,
                                                           double u01x, double u01y, double u10x, double u10y,
                                                           double u11x, double u11y)
{
    using T = typename Elem<kC64>::T;
    using R = typename Elem<kC64>::R;
    T *p = static_cast<T *>(state);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < npairs; k += stride)
    {
        const uint64_t i0 = insert0(k, pivot);
        const uint64_t i1 = i0 ^ delta;
        const bool s0 = sel_eval(packed, i0);
        const uint64_t ie = s0 ? i1 : i0;
        const uint64_t io = s0 ? i0 : i1;
        const T a = p[ie], b = p[io];
        T na, nb;
        na.x = R(u00x) * a.x - R(u00y) * a.y + R(u01x) * b.x - R(u01y) * b.y;
        na.y = R(u00x) * a.y + R(u00y) * a.x + R(u01x) * b.y + R(u01y) * b.x;
        nb.x = R(u10x) * a.x - R(u10y) * a.y + R(u11x) * b.x - R(u11y) * b.y;
        nb.y = R(u10x) * a.y + R(u10y) * a.x + R(u11x) * b.y + R(u11y) * b.x;
        p[ie] = na;
        p[io] = nb;
    }
}

// X / CX explicit permutations (x_block_kernel_cuda / cx_block_kernel_cuda,
// kernel_cuda.cu:442-477): swap i (S_q(i) = 0, and S_cq(i) = 1 for CX) with
// i ^ delta.
template <bool kC64>
__global__ __launch_bounds__(kThreads) void swap_kernel(void *__restrict__ state, uint64_t npairs, uint64_t delta,
                                                        unsigned pivot, uint64_t sel_q, uint64_t sel_cq, int has_ctrl)
{
    using T = typename Elem<kC64>::T;
    T *p = static_cast<T *>(state);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < npairs; k += stride)
    {
        const uint64_t i0 = insert0(k, pivot);
        const uint64_t i1 = i0 ^ delta;
        const uint64_t ie = sel_eval(sel_q, i0) ? i1 : i0;
        if (has_ctrl && !sel_eval(sel_cq, ie))
            continue;
        const T a = p[ie];
        p[ie] = p[ie ^ delta];
        p[ie ^ delta] = a;
    }
}

// Multi-controlled X: swap i0 <-> i0 | 1<<target for every i0 with all
// control bits set and the target bit clear.
template <bool kC64>
__global__ __launch_bounds__(kThreads) void mcx_kernel(void *__restrict__ state, uint64_t npairs,
                                                       const unsigned *__restrict__ fixed_pos, int nfixed,
                                                       uint64_t ctrl_mask, uint64_t tmask)
{
    using T = typename Elem<kC64>::T;
    T *p = static_cast<T *>(state);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < npairs; k += stride)
    {
        uint64_t i0 = k;
        for (int f = 0; f < nfixed; f++)
            i0 = insert0(i0, fixed_pos[f]);
        i0 |= ctrl_mask;
        const uint64_t i1 = i0 | tmask;
        const T a = p[i0];
        p[i0] = p[i1];
        p[i1] = a;
    }
}

// cp_kernel_cuda (kernel_cuda.cu:492-503)
template <bool kC64>
__global__ __launch_bounds__(kThreads) void cp_kernel(void *__restrict__ state, uint64_t nelems, uint64_t sel_cq,
                                                      uint64_t sel_q, double ph_x, double ph_y)
{
    using T = typename Elem<kC64>::T;
    using R = typename Elem<kC64>::R;
    T *p = static_cast<T *>(state);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; x < nelems; x += stride)
    {
        if (!sel_eval(sel_cq, x) || !sel_eval(sel_q, x))
            continue;
        const T a = p[x];
        T r;
        r.x = a.x * R(ph_x) - a.y * R(ph_y);
        r.y = a.x * R(ph_y) + a.y * R(ph_x);
        p[x] = r;
    }
}

template <bool kC64>
__global__ void set_one_kernel(void *state)
{
    using T = typename Elem<kC64>::T;
    T one;
    one.x = 1;
    one.y = 0;
    static_cast<T *>(state)[0] = one;
}

unsigned grid_for(uint64_t items)
{
    const uint64_t blocks = (items + kThreads - 1) / kThreads;
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 16)));
}

template <typename K>
int tile_grid(Ctx &ctx, K kernel, uint64_t tiles)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const uint64_t cap = uint64_t(per_cu) * uint64_t(ctx.num_cus > 0 ? ctx.num_cus : 256);
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(tiles, cap)));
}

template <bool kC64, bool kL1, bool kL2, bool kDiag>
int launch_tile_t(Ctx &ctx, void *state, const TilePass &P, const double *dtab, const double *dcoef)
{
    auto k = u3_tile_kernel<kC64, kL1, kL2, kDiag>;
    hipLaunchKernelGGL(k, dim3(tile_grid(ctx, k, P.ntiles)), dim3(kThreads), 0, ctx.stream, state, P, dtab, dcoef);
    return hip_check(hipGetLastError(), "u3_tile_kernel launch");
}

template <bool kC64, bool kDiag>
int launch_tile_d(Ctx &ctx, void *state, const TilePass &P, bool l1, bool l2, const double *dtab, const double *dcoef)
{
    if (l1 && l2)
        return launch_tile_t<kC64, true, true, kDiag>(ctx, state, P, dtab, dcoef);
    if (l1)
        return launch_tile_t<kC64, true, false, kDiag>(ctx, state, P, dtab, dcoef);
    if (l2)
        return launch_tile_t<kC64, false, true, kDiag>(ctx, state, P, dtab, dcoef);
    return launch_tile_t<kC64, false, false, kDiag>(ctx, state, P, dtab, dcoef);
}

int launch_tile(Ctx &ctx, int prec, void *state, const TilePass &P, bool l1, bool l2, const double *dtab,
                const double *dcoef)
{
    const bool diag = dtab != nullptr;
    if (prec == C64)
        return diag ? launch_tile_d<true, true>(ctx, state, P, l1, l2, dtab, dcoef)
                    : launch_tile_d<true, false>(ctx, state, P, l1, l2, dtab, dcoef);
    return diag ? launch_tile_d<false, true>(ctx, state, P, l1, l2, dtab, dcoef)
                : launch_tile_d<false, false>(ctx, state, P, l1, l2, dtab, dcoef);
}

int layout_of(int prec, int tile_idx)
{
    static const int r128[3][4] = {{8, 9, 10, 11}, {4, 5, 6, 7}, {0, 1, 2, 3}};
    static const int r64[3][4] = {{0, 9, 10, 11}, {1, 2, 3, 4}, {5, 6, 7, 8}};
    const int(*R)[4] = prec == C64 ? r64 : r128;
    for (int l = 0; l < 3; l++)
        for (int i = 0; i < 4; i++)
            if (R[l][i] == tile_idx)
                return l;
    return -1;
}

struct SingleGate
{
    unsigned target;
    const U3Gate *g;
};

// One tile pass over the given targets (at most 4 below bit 4 and 8 above).
int run_tile_pass(Ctx &ctx, void *state, int prec, unsigned n, const std::vector<SingleGate> &gs,
                  const DiagGate *diag = nullptr, size_t ndiag = 0)
{
    std::vector<unsigned> pos = {0, 1, 2, 3};
    std::vector<bool> is_target(64, false);
    for (const SingleGate &s : gs)
        is_target[s.target] = true;
    for (const SingleGate &s : gs)
        if (s.target >= 4)
            pos.push_back(s.target);
    for (unsigned b = 4; b < n && pos.size() < size_t(kTileBits); b++)
        if (!is_target[b])
            pos.push_back(b);
    std::sort(pos.begin(), pos.end());
    TilePass P;
    std::memset(&P, 0, sizeof(P));
    for (int k = 0; k < kTileBits; k++)
        P.tb[k] = pos[k];
    P.ntiles = uint64_t(1) << (n - kTileBits);
    bool use[3] = {false, false, false};
    for (const SingleGate &s : gs)
    {
        const int idx = int(std::find(pos.begin(), pos.end(), s.target) - pos.begin());
        TileGate &tg = P.g[idx];
        const uint64_t alpha = s.g->packed_sel & ((uint64_t(1) << 63) - 1);
        const uint32_t beta = uint32_t(s.g->packed_sel >> 63);
        tg.present = 1;
        tg.alpha_rest = alpha & ~s.g->delta;
        tg.beta = beta;
        if (tg.alpha_rest == 0 && beta)
        {
            // even side is always the target-bit-one element: swap roles
            const double *u = s.g->u;
            const double sw[8] = {u[6], u[7], u[4], u[5], u[2], u[3], u[0], u[1]};
            std::memcpy(tg.u, sw, sizeof(sw));
            tg.beta = 0;
        }
        else
        {
            std::memcpy(tg.u, s.g->u, sizeof(tg.u));
        }
        use[layout_of(prec, idx)] = true;
    }
    const double *dtab = nullptr;
    const double *dcoef = nullptr;
    if (ndiag > 0)
    {
        // phase polynomial of the window, split into tile bits and the rest
        Poly poly;
        window_polynomial(n, diag, ndiag, poly);
        std::vector<double> tab(kTileBits + kTileBits * kTileBits, 0.0);
        int inner[kTileBits];
        std::vector<int> outer;
        for (int i = 0; i < kTileBits; i++)
        {
            inner[i] = int(pos[i]);
            tab[i] = poly.L[pos[i]];
            for (int j = i + 1; j < kTileBits; j++)
                tab[kTileBits + i * kTileBits + j] = poly.q(int(pos[i]), int(pos[j]));
        }
        for (unsigned b = 0; b < n; b++)
            if (std::find(pos.begin(), pos.end(), b) == pos.end())
                outer.push_back(int(b));
        void *d = nullptr;
        ZXA_TRY(ctx.ring.stage(tab.data(), tab.size() * sizeof(double), ctx.stream, &d));
        dtab = static_cast<const double *>(d);
        ZXA_TRY(ctx.scratch_a.ensure(P.ntiles * 16 * sizeof(double)));
        dcoef = static_cast<const double *>(ctx.scratch_a.ptr);
        ZXA_TRY(launch_block_coef(ctx, poly, inner, kTileBits, outer.data(), int(outer.size()),
                                  static_cast<double *>(ctx.scratch_a.ptr)));
    }
    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    ZXA_TRY(launch_tile(ctx, prec, state, P, use[1], use[2], dtab, dcoef));
    ZXA_TRY(ctx.toc(ndiag > 0 ? F_FUSED : F_U3TILE, bytes));
    return ctx.ring.done(ctx.stream);
}

} // namespace

int u3_pair_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate &g)
{
    if (g.delta == 0 || (g.delta >> n) != 0)
    {
        set_error("u3 pair: delta outside the state");
        return EINVAL_;
    }
    const uint64_t npairs = (uint64_t(1) << n) >> 1;
    const unsigned pivot = __builtin_ctzll(g.delta);
    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(u3_pair_kernel<true>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           g.delta, pivot, g.packed_sel, g.u[0], g.u[1], g.u[2], g.u[3], g.u[4], g.u[5], g.u[6],
                           g.u[7]);
    else
        hipLaunchKernelGGL(u3_pair_kernel<false>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state,
                           npairs, g.delta, pivot, g.packed_sel, g.u[0], g.u[1], g.u[2], g.u[3], g.u[4], g.u[5],
                           g.u[6], g.u[7]);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_U3PAIR, bytes);
}

int u3_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate *gates, size_t ngates)
{
    std::vector<SingleGate> singles;
    for (size_t i = 0; i < ngates; i++)
    {
        const U3Gate &g = gates[i];
        const uint64_t alpha = g.packed_sel & ((uint64_t(1) << 63) - 1);
        const bool one_bit = g.delta != 0 && (g.delta & (g.delta - 1)) == 0 && (g.delta >> n) == 0;
        const bool paired = (__builtin_popcountll(alpha & g.delta) & 1) == 1;
        if (n >= unsigned(kTileBits) && one_bit && paired)
            singles.push_back({unsigned(__builtin_ctzll(g.delta)), &g});
        else
            ZXA_TRY(u3_pair_apply(ctx, state, prec, n, g));
    }
    if (singles.empty())
        return OK;
    // targets are distinct (one fused matrix per qubit); group them into passes
    std::sort(singles.begin(), singles.end(),
              [](const SingleGate &a, const SingleGate &b) { return a.target < b.target; });
    std::vector<SingleGate> low, high;
    for (const SingleGate &s : singles)
        (s.target < 4 ? low : high).push_back(s);
    for (size_t i = 1; i < singles.size(); i++)
        if (singles[i].target == singles[i - 1].target)
        {
            set_error("u3_apply: two gates on one target in one batch");
            return EINVAL_;
        }
    size_t hi = 0;
    bool first = true;
    while (first || hi < high.size())
    {
        std::vector<SingleGate> pass;
        if (first)
            pass = low;
        const size_t take = std::min<size_t>(8, high.size() - hi);
        for (size_t k = 0; k < take; k++)
            pass.push_back(high[hi + k]);
        hi += take;
        first = false;
        if (!pass.empty())
            ZXA_TRY(run_tile_pass(ctx, state, prec, n, pass));
    }
    return OK;
}

int fused_pass_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate *gates, size_t ngates,
                     const DiagGate *diag, size_t ndiag)
{
    if (n < unsigned(kTileBits))
    {
        set_error("fused pass: state narrower than a tile");
        return EINVAL_;
    }
    if (ndiag > 0 && !window_is_polynomial(n, diag, ndiag))
    {
        set_error("fused pass: diagonal window is not a phase polynomial");
        return EINVAL_;
    }
    std::vector<SingleGate> gs;
    int low = 0, high = 0;
    for (size_t i = 0; i < ngates; i++)
    {
        const U3Gate &g = gates[i];
        const uint64_t alpha = g.packed_sel & ((uint64_t(1) << 63) - 1);
        if (g.delta == 0 || (g.delta & (g.delta - 1)) != 0 || alpha != g.delta || (g.delta >> n) != 0)
        {
            set_error("fused pass: gate is not a single-bit U3");
            return EINVAL_;
        }
        const unsigned t = unsigned(__builtin_ctzll(g.delta));
        for (const SingleGate &o : gs)
            if (o.target == t)
            {
                set_error("fused pass: two gates on one target");
                return EINVAL_;
            }
        (t < 4 ? low : high)++;
        gs.push_back({t, &g});
    }
    if (high > 8)
    {
        set_error("fused pass: more than 8 targets above bit 3");
        return EINVAL_;
    }
    if (gs.empty())
        return diag_apply(ctx, state, prec, n, diag, ndiag);
    return run_tile_pass(ctx, state, prec, n, gs, diag, ndiag);
}

int x_apply(Ctx &ctx, void *state, int prec, unsigned n, uint64_t delta, uint64_t sel)
{
    if (delta == 0 || (delta >> n) != 0)
    {
        set_error("x: delta outside the state");
        return EINVAL_;
    }
    const uint64_t npairs = (uint64_t(1) << n) >> 1;
    const unsigned pivot = __builtin_ctzll(delta);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(swap_kernel<true>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           delta, pivot, sel, uint64_t(0), 0);
    else
        hipLaunchKernelGGL(swap_kernel<false>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           delta, pivot, sel, uint64_t(0), 0);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_PERM, 2ull * (uint64_t(1) << n) * elem_bytes(prec));
}

int cx_apply(Ctx &ctx, void *state, int prec, unsigned n, uint64_t delta, uint64_t sel_cq, uint64_t sel_q)
{
    if (delta == 0 || (delta >> n) != 0)
    {
        set_error("cx: delta outside the state");
        return EINVAL_;
    }
    const uint64_t npairs = (uint64_t(1) << n) >> 1;
    const unsigned pivot = __builtin_ctzll(delta);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(swap_kernel<true>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           delta, pivot, sel_q, sel_cq, 1);
    else
        hipLaunchKernelGGL(swap_kernel<false>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           delta, pivot, sel_q, sel_cq, 1);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_PERM, 2ull * (uint64_t(1) << n) * elem_bytes(prec));
}

int mcx_apply(Ctx &ctx, void *state, int prec, unsigned n, uint64_t ctrl_mask, unsigned target)
{
    if (target >= n || (ctrl_mask >> n) != 0 || ((ctrl_mask >> target) & 1u))
    {
        set_error("mcx: target/controls outside the state or overlapping");
        return EINVAL_;
    }
    std::vector<unsigned> fixed;
    for (unsigned b = 0; b < n; b++)
        if (((ctrl_mask >> b) & 1u) || b == target)
            fixed.push_back(b);
    const uint64_t npairs = uint64_t(1) << (n - fixed.size());
    void *dpos = nullptr;
    ZXA_TRY(ctx.ring.stage(fixed.data(), fixed.size() * sizeof(unsigned), ctx.stream, &dpos));
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(mcx_kernel<true>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           static_cast<const unsigned *>(dpos), int(fixed.size()), ctrl_mask, uint64_t(1) << target);
    else
        hipLaunchKernelGGL(mcx_kernel<false>, dim3(grid_for(npairs)), dim3(kThreads), 0, ctx.stream, state, npairs,
                           static_cast<const unsigned *>(dpos), int(fixed.size()), ctrl_mask, uint64_t(1) << target);
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_PERM, 4ull * npairs * elem_bytes(prec)));
    return ctx.ring.done(ctx.stream);
}

int cp_apply(Ctx &ctx, void *state, int prec, unsigned n, uint64_t sel_cq, uint64_t sel_q, double theta)
{
    const uint64_t nelems = uint64_t(1) << n;
    const double c = std::cos(theta), s = std::sin(theta);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(cp_kernel<true>, dim3(grid_for(nelems)), dim3(kThreads), 0, ctx.stream, state, nelems,
                           sel_cq, sel_q, c, s);
    else
        hipLaunchKernelGGL(cp_kernel<false>, dim3(grid_for(nelems)), dim3(kThreads), 0, ctx.stream, state, nelems,
                           sel_cq, sel_q, c, s);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_CP, 2ull * nelems * elem_bytes(prec));
}

int state_reset(Ctx &ctx, void *state, int prec, size_t alloc_elems)
{
    ZXA_TRY(ctx.tic());
    ZXA_HIP(hipMemsetAsync(state, 0, alloc_elems * elem_bytes(prec), ctx.stream));
    if (prec == C64)
        hipLaunchKernelGGL(set_one_kernel<true>, dim3(1), dim3(1), 0, ctx.stream, state);
    else
        hipLaunchKernelGGL(set_one_kernel<false>, dim3(1), dim3(1), 0, ctx.stream, state);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_RESET, uint64_t(alloc_elems) * elem_bytes(prec));
}

} // namespace zxa
