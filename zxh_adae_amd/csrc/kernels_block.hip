// kernels_block.hip -- dense 5-qubit blocks (32x32 complex unitaries) as
// matrix-core products.  No reference counterpart: the reference applies only
// single-qubit U3 gates and diagonal windows.
//
// A block on target qubits t0..t4 maps every group of 32 amplitudes that
// differ only in those bits, y[t] = sum_t' U[t][t'] x[t'] (bit b of the
// combination index t is qubit t_b).  In real form a group is a 64-vector
// (32 real parts, then 32 imaginary parts) and U the 64x64 matrix
// M = [[Ur, -Ui], [Ui, Ur]], so a wave holding G groups computes
// Y (64 x G) = M X (64 x G) on the matrix cores:
//   complex64:  v_mfma_f32_32x32x2f32, G = 32 groups (1024 amplitudes), two
//               32-row output blocks, 32 k-steps of 2 (FP32 accumulation);
//   complex128: v_mfma_f64_16x16x4f64, G = 16 groups (512 amplitudes), four
//               16-row output blocks, 16 k-steps of 4 (FP64).
// Operand layouts (measured on gfx950 with a one-wave probe): A[i][k] and
// B[k][j] in lane i + G k resp. j + G k of the step; D of the f32 32x32x2
// holds row 8 (r / 4) + 4 (lane / 32) + r % 4 in register r, D of the f64
// 16x16x4 holds row lane / 16 + 4 r, both in column lane % G.
// The group index is the lane (mod G), so the X operand is loaded straight
// from HBM into the MFMA operand layout: lane (n, h) holds the amplitudes of
// group n whose combination index t is congruent to h mod 64/G; the state is
// not staged through LDS.  The group bits are the lowest non-target address
// bits, so a load instruction reads runs of G consecutive amplitudes (256 B)
// when no target is among the lowest bits.  When targets are low bits the
// runs of one instruction shrink to 8 or 16 bytes, and the wave's 16 (8) load
// instructions together cover whole lines: ordinary cached loads and stores
// then let L2 merge them, where nontemporal ones fetch each line several
// times (a c64 block on qubits 0, 1, 7, 15, 22 of 33: 84 ms nontemporal,
// 33 ms cached).  Variants (knob block.variant.<prec>): prefetch of the next
// wave tile, the matrix re-read from LDS per tile (fewer registers, more
// waves) or held in registers, cached or nontemporal access; and (knob
// block.tiles.<prec>) a persistent grid or a fixed number of wave tiles per
// wave.  The defaults were chosen on the 32/33-qubit random circuits:
// complex64 prefetch + LDS + cached (28.2 ms per 33-qubit sweep), complex128
// registers + cached (27.1 ms per 32-qubit sweep; prefetching with FP64
// matrix cores busy was slower), both on persistent grids (four wave tiles
// per wave gave 28.4 and 27.3 ms over the whole circuits).
// M is pre-arranged on the host in the A operand's lane order and read from
// LDS, one word per lane per step (conflict-free).  Each wave owns disjoint
// groups: it reads them all, multiplies, and writes them back.
#include "dev.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

namespace zxa
{
namespace
{

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

typedef float fx16 __attribute__((ext_vector_type(16)));
typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
typedef double dvec2 __attribute__((ext_vector_type(2)));

#define ZXA_CONST const __attribute__((address_space(4)))

struct BArgs
{
    uint32_t sorted[10]; // target and group bits, ascending
    uint32_t nbits;      // 10 (complex64) or 9 (complex128)
    uint32_t pad;
    uint64_t nwtiles; // wave tiles (G groups each)
    uint64_t gstep;   // total waves of the grid, deposited into the free bits
    uint64_t tmask;   // mask of the target and group bits
};

// Uniform target-bit offsets: tin[s] for input step s, tout[o] for output o
// (the lane adds its group and h offsets: lane_off[lane], lane_off[64+lane]).
struct BTab
{
    uint64_t tin[16];
    uint64_t tout[16];
};

__host__ __device__ __forceinline__ uint64_t deposit_bits(uint64_t w, const uint32_t *sorted, unsigned nb)
{
    for (unsigned k = 0; k < nb; k++)
    {
        const unsigned p = sorted[k];
        w = ((w >> p) << (p + 1)) | (w & ((uint64_t(1) << p) - 1));
    }
    return w;
}

// The matrix is re-read from LDS at every tile: an opaque zero added to the
// LDS index keeps the compiler from hoisting the 64 loop-invariant operand
// words per lane into registers, which would cost 64 (complex64) or 128
// (complex128) VGPRs and halve the waves per SIMD.
template <bool kNT, typename V>
__device__ __forceinline__ V ld(const V *p)
{
    if constexpr (kNT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <bool kNT, typename V>
__device__ __forceinline__ void st(V v, V *p)
{
    if constexpr (kNT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <bool kLds>
__device__ __forceinline__ int opaque_zero()
{
    int z = 0;
    if constexpr (kLds)
        asm volatile("" : "+v"(z));
    return z;
}

// One wave tile of complex64: Y = M X on the 32 groups in x, stored at base.
template <bool kLds, bool kNT>
__device__ __forceinline__ void block_c64_tile(const fvec2 (&x)[16], const float *fm, int lane, fvec2 *state,
                                               uint64_t at, ZXA_CONST BTab *T)
{
    const int z = opaque_zero<kLds>();
    fx16 c0 = {}, c1 = {};
#pragma unroll
    for (int s = 0; s < 32; s++)
    {
        const float b = s < 16 ? x[s].x : x[s - 16].y;
        c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(fm[(s * 2 + 0) * 64 + lane + z], b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(fm[(s * 2 + 1) * 64 + lane + z], b, c1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++)
        st<kNT>(fvec2{c0[r], c1[r]}, state + (at | T->tout[r]));
}

// complex64: frag[(s * 2 + blk) * 64 + lane] = M[32 blk + lane % 32][2 s + lane / 32]
// Each wave streams through its wave tiles with two register buffers: the
// loads of the next tile are issued before the current tile's products, so
// HBM stays busy while the matrix cores run (past the last tile the prefetch
// re-reads the current one).
template <bool kPre, bool kLds, bool kNT>
__global__ __launch_bounds__(kThreads) void block_kernel_c64(fvec2 *__restrict__ state, const BArgs A,
                                                             const BTab *__restrict__ tabp,
                                                             const float *__restrict__ frag,
                                                             const uint64_t *__restrict__ lane_off)
{
    __shared__ float fm[32 * 2 * 64];
    for (int i = threadIdx.x; i < 32 * 2 * 64; i += kThreads)
        fm[i] = frag[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t gin = lane_off[lane];
    const uint64_t gout = lane_off[64 + lane];
    ZXA_CONST BTab *T = (ZXA_CONST BTab *)tabp;
    uint64_t w = uint64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    const uint64_t wstride = uint64_t(gridDim.x) * kWaves;
    if (w >= A.nwtiles)
        return;
    uint64_t b0 = deposit_bits(w, A.sorted, A.nbits);
    fvec2 xa[16], xb[16];
#pragma unroll
    for (int s = 0; s < 16; s++)
        xa[s] = ld<kNT>(state + (b0 | gin | T->tin[s]));
    if constexpr (!kPre)
    {
        for (;;)
        {
            block_c64_tile<kLds, kNT>(xa, fm, lane, state, b0 | gout, T);
            w += wstride;
            if (w >= A.nwtiles)
                return;
            b0 = ((b0 | A.tmask) + A.gstep) & ~A.tmask;
#pragma unroll
            for (int s = 0; s < 16; s++)
                xa[s] = ld<kNT>(state + (b0 | gin | T->tin[s]));
        }
    }
    for (;;)
    {
        const bool m1 = w + wstride < A.nwtiles;
        const uint64_t b1 = m1 ? ((b0 | A.tmask) + A.gstep) & ~A.tmask : b0;
#pragma unroll
        for (int s = 0; s < 16; s++)
            xb[s] = ld<kNT>(state + (b1 | gin | T->tin[s]));
        block_c64_tile<kLds, kNT>(xa, fm, lane, state, b0 | gout, T);
        if (!m1)
            break;
        w += wstride;
        const bool m2 = w + wstride < A.nwtiles;
        const uint64_t b2 = m2 ? ((b1 | A.tmask) + A.gstep) & ~A.tmask : b1;
#pragma unroll
        for (int s = 0; s < 16; s++)
            xa[s] = ld<kNT>(state + (b2 | gin | T->tin[s]));
        block_c64_tile<kLds, kNT>(xb, fm, lane, state, b1 | gout, T);
        if (!m2)
            break;
        w += wstride;
        b0 = b2;
    }
}

// One wave tile of complex128.
template <bool kLds, bool kNT>
__device__ __forceinline__ void block_c128_tile(const dvec2 (&x)[8], const double *fm, int lane, dvec2 *state,
                                                uint64_t at, ZXA_CONST BTab *T)
{
    const int z = opaque_zero<kLds>();
    dx4 c[4] = {};
#pragma unroll
    for (int s = 0; s < 16; s++)
    {
        const double b = s < 8 ? x[s].x : x[s - 8].y;
#pragma unroll
        for (int blk = 0; blk < 4; blk++)
            c[blk] = __builtin_amdgcn_mfma_f64_16x16x4f64(fm[(s * 4 + blk) * 64 + lane + z], b, c[blk], 0, 0, 0);
    }
#pragma unroll
    for (int o = 0; o < 8; o++)
    {
        const int bp = o >> 2, r = o & 3;
        st<kNT>(dvec2{c[bp][r], c[bp + 2][r]}, state + (at | T->tout[o]));
    }
}

// complex128: frag[(s * 4 + blk) * 64 + lane] = M[16 blk + lane % 16][4 s + lane / 16]
template <bool kPre, bool kLds, bool kNT>
__global__ __launch_bounds__(kThreads) void block_kernel_c128(dvec2 *__restrict__ state, const BArgs A,
                                                              const BTab *__restrict__ tabp,
                                                              const double *__restrict__ frag,
                                                              const uint64_t *__restrict__ lane_off)
{
    __shared__ double fm[16 * 4 * 64];
    for (int i = threadIdx.x; i < 16 * 4 * 64; i += kThreads)
        fm[i] = frag[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t gin = lane_off[lane];
    const uint64_t gout = lane_off[64 + lane];
    ZXA_CONST BTab *T = (ZXA_CONST BTab *)tabp;
    uint64_t w = uint64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    const uint64_t wstride = uint64_t(gridDim.x) * kWaves;
    if (w >= A.nwtiles)
        return;
    uint64_t b0 = deposit_bits(w, A.sorted, A.nbits);
    dvec2 xa[8], xb[8];
#pragma unroll
    for (int s = 0; s < 8; s++)
        xa[s] = ld<kNT>(state + (b0 | gin | T->tin[s]));
    if constexpr (!kPre)
    {
        for (;;)
        {
            block_c128_tile<kLds, kNT>(xa, fm, lane, state, b0 | gout, T);
            w += wstride;
            if (w >= A.nwtiles)
                return;
            b0 = ((b0 | A.tmask) + A.gstep) & ~A.tmask;
#pragma unroll
            for (int s = 0; s < 8; s++)
                xa[s] = ld<kNT>(state + (b0 | gin | T->tin[s]));
        }
    }
    for (;;)
    {
        const bool m1 = w + wstride < A.nwtiles;
        const uint64_t b1 = m1 ? ((b0 | A.tmask) + A.gstep) & ~A.tmask : b0;
#pragma unroll
        for (int s = 0; s < 8; s++)
            xb[s] = ld<kNT>(state + (b1 | gin | T->tin[s]));
        block_c128_tile<kLds, kNT>(xa, fm, lane, state, b0 | gout, T);
        if (!m1)
            break;
        w += wstride;
        const bool m2 = w + wstride < A.nwtiles;
        const uint64_t b2 = m2 ? ((b1 | A.tmask) + A.gstep) & ~A.tmask : b1;
#pragma unroll
        for (int s = 0; s < 8; s++)
            xa[s] = ld<kNT>(state + (b2 | gin | T->tin[s]));
        block_c128_tile<kLds, kNT>(xb, fm, lane, state, b1 | gout, T);
        if (!m2)
            break;
        w += wstride;
        b0 = b2;
    }
}

template <typename K>
int grid_for_waves(Ctx &ctx, K kernel, uint64_t nwtiles)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const uint64_t cap = uint64_t(per_cu) * uint64_t(ctx.num_cus > 0 ? ctx.num_cus : 256);
    const uint64_t need = (nwtiles + kWaves - 1) / kWaves;
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(need, cap)));
}

} // namespace

int block5_apply(Ctx &ctx, void *state, int prec, unsigned n, const unsigned *targets, const double *u)
{
    const bool c64 = prec == C64;
    const unsigned G = c64 ? 32u : 16u;          // groups per wave
    const unsigned gbits = c64 ? 5u : 4u;        // group address bits
    const unsigned wbits = 5u + gbits;           // wave-tile bits
    if (n < wbits || n > 40)
    {
        set_error("block5: state narrower than one wave tile (10 bits complex64, 9 complex128) or wider than 40");
        return EINVAL_;
    }
    uint64_t tm = 0;
    for (int b = 0; b < 5; b++)
    {
        if (targets[b] >= n || ((tm >> targets[b]) & 1))
        {
            set_error("block5: targets must be 5 distinct qubits of the state");
            return EINVAL_;
        }
        tm |= uint64_t(1) << targets[b];
    }
    // group bits: the lowest non-target bits
    std::vector<unsigned> gpos;
    for (unsigned b = 0; b < n && gpos.size() < gbits; b++)
        if (!((tm >> b) & 1))
            gpos.push_back(b);
    BArgs A;
    std::memset(&A, 0, sizeof(A));
    std::vector<unsigned> all(targets, targets + 5);
    all.insert(all.end(), gpos.begin(), gpos.end());
    std::sort(all.begin(), all.end());
    for (unsigned k = 0; k < wbits; k++)
    {
        A.sorted[k] = all[k];
        A.tmask |= uint64_t(1) << all[k];
    }
    A.nbits = wbits;
    A.nwtiles = uint64_t(1) << (n - wbits);
    auto toff = [&](unsigned t) {
        uint64_t o = 0;
        for (int b = 0; b < 5; b++)
            if ((t >> b) & 1)
                o |= uint64_t(1) << targets[b];
        return o;
    };
    auto goff = [&](unsigned g) {
        uint64_t o = 0;
        for (unsigned b = 0; b < gbits; b++)
            if ((g >> b) & 1)
                o |= uint64_t(1) << gpos[b];
        return o;
    };
    BTab T;
    std::memset(&T, 0, sizeof(T));
    std::vector<uint64_t> lane_off(128);
    std::vector<double> M(64 * 64);
    for (int i = 0; i < 32; i++)
        for (int k = 0; k < 32; k++)
        {
            const double re = u[2 * (i * 32 + k)], im = u[2 * (i * 32 + k) + 1];
            M[i * 64 + k] = re;
            M[i * 64 + 32 + k] = -im;
            M[(32 + i) * 64 + k] = im;
            M[(32 + i) * 64 + 32 + k] = re;
        }
    std::vector<float> f32;
    std::vector<double> f64;
    if (c64)
    {
        // input t = 2 s + h; output t = 8 (r / 4) + 4 h + r % 4
        for (int s = 0; s < 16; s++)
            T.tin[s] = toff(2 * s);
        for (int r = 0; r < 16; r++)
            T.tout[r] = toff(8 * (r / 4) + (r % 4));
        for (int l = 0; l < 64; l++)
        {
            const unsigned nn = l % 32, h = l / 32;
            lane_off[l] = goff(nn) | toff(h);
            lane_off[64 + l] = goff(nn) | toff(4 * h);
        }
        f32.resize(32 * 2 * 64);
        for (int s = 0; s < 32; s++)
            for (int blk = 0; blk < 2; blk++)
                for (int l = 0; l < 64; l++)
                    f32[(s * 2 + blk) * 64 + l] = float(M[(32 * blk + l % 32) * 64 + 2 * s + l / 32]);
    }
    else
    {
        // input t = 4 s + h; output (D register r of block b' holds row
        // h + 4 r) t = 16 b' + 4 r + h (o = 4 b' + r)
        for (int s = 0; s < 8; s++)
            T.tin[s] = toff(4 * s);
        for (int o = 0; o < 8; o++)
            T.tout[o] = toff(16 * (o / 4) + 4 * (o % 4));
        for (int l = 0; l < 64; l++)
        {
            const unsigned nn = l % 16, h = l / 16;
            lane_off[l] = goff(nn) | toff(h);
            lane_off[64 + l] = goff(nn) | toff(h);
        }
        f64.resize(16 * 4 * 64);
        for (int s = 0; s < 16; s++)
            for (int blk = 0; blk < 4; blk++)
                for (int l = 0; l < 64; l++)
                    f64[(s * 4 + blk) * 64 + l] = M[(16 * blk + l % 16) * 64 + 4 * s + l / 16];
    }
    (void)G;
    void *dT = nullptr, *dL = nullptr, *dF = nullptr;
    ZXA_TRY(ctx.ring.stage(&T, sizeof(T), ctx.stream, &dT));
    ZXA_TRY(ctx.ring.stage(lane_off.data(), lane_off.size() * sizeof(uint64_t), ctx.stream, &dL));
    if (c64)
        ZXA_TRY(ctx.ring.stage(f32.data(), f32.size() * sizeof(float), ctx.stream, &dF));
    else
        ZXA_TRY(ctx.ring.stage(f64.data(), f64.size() * sizeof(double), ctx.stream, &dF));
    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    // kernel variant (knob block.variant.<prec>): bit 0 prefetches the next
    // wave tile, bit 1 re-reads the matrix from LDS at every tile, bit 2
    // uses ordinary (cached) loads and stores instead of nontemporal ones
    const int var = tune_get(c64 ? "block.variant.c64" : "block.variant.c128");
    // wave tiles per wave (knob block.tiles.<prec>): 0 = persistent grid
    // (occupancy x CUs workgroups), else that many tiles per wave
    const int per_wave = tune_get(c64 ? "block.tiles.c64" : "block.tiles.c128");
    auto launch = [&](auto kernel, auto *st, const auto *frag) {
        int grid = grid_for_waves(ctx, kernel, A.nwtiles);
        if (per_wave > 0)
        {
            const uint64_t per_wg = uint64_t(per_wave) * kWaves;
            grid = static_cast<int>(std::min<uint64_t>((A.nwtiles + per_wg - 1) / per_wg, uint64_t(INT_MAX)));
        }
        A.gstep = deposit_bits(uint64_t(grid) * kWaves, A.sorted, A.nbits);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, ctx.stream, st, A, static_cast<const BTab *>(dT),
                           frag, static_cast<const uint64_t *>(dL));
    };
    if (c64)
    {
        fvec2 *sp = static_cast<fvec2 *>(state);
        const float *fr = static_cast<const float *>(dF);
        switch (var & 7)
        {
        case 0: launch(block_kernel_c64<false, false, true>, sp, fr); break;
        case 1: launch(block_kernel_c64<true, false, true>, sp, fr); break;
        case 2: launch(block_kernel_c64<false, true, true>, sp, fr); break;
        case 3: launch(block_kernel_c64<true, true, true>, sp, fr); break;
        case 4: launch(block_kernel_c64<false, false, false>, sp, fr); break;
        case 5: launch(block_kernel_c64<true, false, false>, sp, fr); break;
        case 6: launch(block_kernel_c64<false, true, false>, sp, fr); break;
        default: launch(block_kernel_c64<true, true, false>, sp, fr); break;
        }
    }
    else
    {
        dvec2 *sp = static_cast<dvec2 *>(state);
        const double *fr = static_cast<const double *>(dF);
        switch (var & 7)
        {
        case 0: launch(block_kernel_c128<false, false, true>, sp, fr); break;
        case 1: launch(block_kernel_c128<true, false, true>, sp, fr); break;
        case 2: launch(block_kernel_c128<false, true, true>, sp, fr); break;
        case 3: launch(block_kernel_c128<true, true, true>, sp, fr); break;
        case 4: launch(block_kernel_c128<false, false, false>, sp, fr); break;
        case 5: launch(block_kernel_c128<true, false, false>, sp, fr); break;
        case 6: launch(block_kernel_c128<false, true, false>, sp, fr); break;
        default: launch(block_kernel_c128<true, true, false>, sp, fr); break;
        }
    }
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_BLOCK, bytes));
    return ctx.ring.done(ctx.stream);
}

} // namespace zxa
