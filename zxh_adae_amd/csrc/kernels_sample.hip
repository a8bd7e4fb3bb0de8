// kernels_sample.hip -- multinomial sampling, bit-exact with the reference's
// streaming CPU sampler (reference src/zxh/src/measure/measure_general.cpp,
// measure_streaming_impl.h, measure_common.cpp).
//
// The reference builds, per segment of 2^I amplitudes, a CDF from a Blelloch
// exclusive scan of p_i = re^2 + im^2, offsets it by a binary-counter prefix
// of earlier segment totals, and resolves sorted thresholds with
// std::upper_bound.  Every value it compares is a node sum of ONE balanced
// binary tree over p (Blelloch up-sweep nodes, merge_prob_sum_range, and the
// binary counter all add left+right children of aligned power-of-two blocks),
// combined left to right in Fenwick order:
//     excl_j   = (((0 + N_{l1}[.]) + N_{l2}[.]) + ...)   set bits of j, high to low
//     prefix_s = same over segment indices, with levels >= I
//     cdf_j    = (prefix_s + excl_j) + p_j,   tail_s = prefix_s + N_I[s]
// So the GPU stores the tree levels >= LS (LS = min(4, I) for complex128,
// min(5, I) for complex64; lower levels are recomputed from the 256 bytes of
// amplitudes of a block: 16 complex128 or 32 complex64), and each shot performs
// libstdc++'s exact upper_bound probe sequence, evaluating cdf at each probe.
// A threshold belongs to the first segment whose running-max tail exceeds
// it, which is where the reference's sequential segment walk places it.
// No fused multiply-add is allowed anywhere on this path.
#include "dev.h"

#include <algorithm>
#include <cstring>
#include <random>
#include <vector>

#pragma clang fp contract(off)

namespace zxa
{
namespace
{

constexpr int kThreads = 256;
constexpr int kSpanBits = 12; // input values per workgroup in the tree kernel
constexpr int kMaxLevels = 64;

struct Levels
{
    uint64_t off[kMaxLevels]; // offset of level l in the tree array (valid for LS <= l <= m)
};

__device__ __forceinline__ double p_of(const void *state, bool c64, uint64_t i)
{
    if (c64)
    {
        const float2 a = static_cast<const float2 *>(state)[i];
        const double x = a.x, y = a.y;
        return __dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y));
    }
    const double2 a = static_cast<const double2 *>(state)[i];
    return __dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y));
}

// One workgroup reduces up to 2^kSpanBits input values of level lin into the
// levels lin+1 .. lin+span_bits and stores every level in [LS, m] it owns.
// src_kind: 0 = c128 amplitudes, 1 = c64 amplitudes, 2 = tree nodes (f64).
__global__ __launch_bounds__(kThreads) void tree_kernel(const void *__restrict__ src, int src_kind, int lin,
                                                        int span_bits, int LS, int m, double *__restrict__ tree,
                                                        const Levels lv)
{
    __shared__ double buf[1 << kSpanBits];
    const int span = 1 << span_bits;
    const uint64_t g0 = static_cast<uint64_t>(blockIdx.x) << span_bits;
    for (int i = threadIdx.x; i < span; i += kThreads)
    {
        double v;
        if (src_kind == 2)
            v = static_cast<const double *>(src)[lv.off[lin] + g0 + i];
        else
            v = p_of(src, src_kind == 1, g0 + i);
        buf[i] = v;
        if (src_kind != 2 && lin >= LS)
            tree[lv.off[lin] + g0 + i] = v;
    }
    __syncthreads();
    for (int r = 1; r <= span_bits; r++)
    {
        const int cnt = span >> r;
        const int l = lin + r;
        double vals[(1 << kSpanBits) / kThreads / 2 > 0 ? (1 << kSpanBits) / kThreads / 2 : 1];
        int nv = 0;
        for (int k = threadIdx.x; k < cnt; k += kThreads)
            vals[nv++] = __dadd_rn(buf[2 * k], buf[2 * k + 1]);
        __syncthreads();
        nv = 0;
        for (int k = threadIdx.x; k < cnt; k += kThreads)
        {
            const double v = vals[nv++];
            buf[k] = v;
            if (l >= LS && l <= m)
                tree[lv.off[l] + (g0 >> r) + k] = v;
        }
        __syncthreads();
    }
}

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

// tree_amp_kernel geometry: a wave tile is 16 loads of 16 bytes per lane,
// 1024 amplitudes (c128) or 2048 (c64); a workgroup holds four tiles.
template <bool kC64>
struct TreeGeo
{
    static constexpr int kTileBits = kC64 ? 11 : 10;
    static constexpr int kRounds = kC64 ? 2 : 1;     // 512 level-1 nodes per round
    static constexpr int kGroupBits = kTileBits + 2; // amplitudes per workgroup
};
constexpr int kRowPad = 9; // LDS doubles per row of 8 level-1 nodes (one pad)

// Orders this wave's LDS accesses (a wave executes in lockstep; the fences
// keep the compiler from moving accesses across the exchange).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The value of the neighbouring lane (lane ^ 1), by DPP quad permutation.
__device__ __forceinline__ double lane_xor1(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp(int(b), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), 0xB1, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// First tree pass when the stored levels start at the precision's low level
// (4 for complex128, 5 for complex64): one read of the state, that level
// written; the upper levels come from it in tree_kernel passes.
// Each wave reduces a tile.  All of the tile's loads (coalesced, 16 bytes per
// lane, nontemporal) are issued before any arithmetic.  Level 1 is formed in
// registers: inside the lane for complex64 (a 16-byte load holds two
// amplitudes), with the neighbouring lane by DPP for complex128.  Level-1
// nodes go to LDS in rows of 8 (16 amplitudes; padded to 9 doubles so that the
// row reads are conflict-free); each lane sums one row as the aligned-pair
// tree (levels 2..4) in registers; complex128 stores the level-4 node (a
// contiguous 512 bytes per wave), complex64 adds the neighbouring lane's (DPP)
// and stores level 5 from the even lanes (256 bytes per wave and round).  The
// stores are nontemporal: the state streams through without the tree's writes
// breaking up the reads.  Every node is
// left + right of the same two children as in the reference's balanced tree
// (IEEE addition is commutative, so the order of the two operands does not
// matter).
template <bool kC64>
__global__ __launch_bounds__(kThreads) void tree_amp_kernel(const void *__restrict__ state, double *__restrict__ tl)
{
    using G = TreeGeo<kC64>;
    __shared__ double rows[kThreads / 64][64 * kRowPad];
    const int w = threadIdx.x >> 6, L = threadIdx.x & 63;
    const uint64_t tile = uint64_t(blockIdx.x) * (kThreads / 64) + w;
    const uint64_t base = tile << G::kTileBits;
    double *rb = rows[w];
    if (kC64)
    {
        const fvec4 *src = static_cast<const fvec4 *>(state) + (base >> 1);
        fvec4 v[16];
#pragma unroll
        for (int k = 0; k < 16; k++)
            v[k] = __builtin_nontemporal_load(src + 64 * k + L);
#pragma unroll
        for (int r = 0; r < 2; r++)
        {
#pragma unroll
            for (int k = 0; k < 8; k++)
            {
                const fvec4 a = v[8 * r + k];
                const double x0 = a.x, y0 = a.y, x1 = a.z, y1 = a.w;
                const double n1 = __dadd_rn(__dadd_rn(__dmul_rn(x0, x0), __dmul_rn(y0, y0)),
                                            __dadd_rn(__dmul_rn(x1, x1), __dmul_rn(y1, y1)));
                const int j = 64 * k + L; // level-1 node within the round
                rb[(j >> 3) * kRowPad + (j & 7)] = n1;
            }
            wave_sync();
            double q[8];
#pragma unroll
            for (int c = 0; c < 8; c++)
                q[c] = rb[L * kRowPad + c];
            wave_sync();
#pragma unroll
            for (int c = 0; c < 4; c++)
                q[c] = __dadd_rn(q[2 * c], q[2 * c + 1]);
            // level 4 of row L, then level 5 with the neighbouring lane
            const double n4 = __dadd_rn(__dadd_rn(q[0], q[1]), __dadd_rn(q[2], q[3]));
            const double n5 = __dadd_rn(n4, lane_xor1(n4));
            if ((L & 1) == 0)
                __builtin_nontemporal_store(n5, tl + (base >> 5) + 32 * r + (L >> 1));
        }
    }
    else
    {
        const dvec2 *src = static_cast<const dvec2 *>(state) + base;
        dvec2 v[16];
#pragma unroll
        for (int k = 0; k < 16; k++)
            v[k] = __builtin_nontemporal_load(src + 64 * k + L);
#pragma unroll
        for (int k = 0; k < 16; k++)
        {
            const double p = __dadd_rn(__dmul_rn(v[k].x, v[k].x), __dmul_rn(v[k].y, v[k].y));
            const double n1 = __dadd_rn(p, lane_xor1(p));
            const int j = 32 * k + (L >> 1); // level-1 node within the tile
            if ((L & 1) == 0)
                rb[(j >> 3) * kRowPad + (j & 7)] = n1;
        }
        wave_sync();
        double q[8];
#pragma unroll
        for (int c = 0; c < 8; c++)
            q[c] = rb[L * kRowPad + c];
#pragma unroll
        for (int c = 0; c < 4; c++)
            q[c] = __dadd_rn(q[2 * c], q[2 * c + 1]);
        __builtin_nontemporal_store(__dadd_rn(__dadd_rn(q[0], q[1]), __dadd_rn(q[2], q[3])), tl + (base >> 4) + L);
    }
}

// Tree pass over stored nodes: level lin -> levels lin+1 .. lin+5, all
// stored.  A wave tile is 2048 level-lin nodes (16 loads of two nodes per
// lane); level lin+1 is the in-lane pair (stored from registers, 512
// contiguous bytes per load), then each lane sums a row of 16 level-(lin+1)
// nodes from LDS (rows padded to 17 doubles) as the aligned-pair tree and
// stores its 8 + 4 + 2 + 1 nodes of levels lin+2 .. lin+5, which are
// contiguous across the wave's lanes.  Needs 2^13 level-lin nodes per
// workgroup.
constexpr int kUpRowPad = 17;
__global__ __launch_bounds__(kThreads) void tree_up_kernel(const double *__restrict__ in, double *__restrict__ o1,
                                                           double *__restrict__ o2, double *__restrict__ o3,
                                                           double *__restrict__ o4, double *__restrict__ o5)
{
    __shared__ double rows[kThreads / 64][64 * kUpRowPad];
    const int w = threadIdx.x >> 6, L = threadIdx.x & 63;
    const uint64_t tile = uint64_t(blockIdx.x) * (kThreads / 64) + w; // 1024 level-(lin+1) nodes each
    const dvec2 *src = reinterpret_cast<const dvec2 *>(in) + (tile << 10);
    double *rb = rows[w];
    dvec2 v[16];
#pragma unroll
    for (int k = 0; k < 16; k++)
        v[k] = __builtin_nontemporal_load(src + 64 * k + L);
#pragma unroll
    for (int k = 0; k < 16; k++)
    {
        const double n = __dadd_rn(v[k].x, v[k].y);
        const int j = 64 * k + L;
        o1[(tile << 10) + j] = n;
        rb[(j >> 4) * kUpRowPad + (j & 15)] = n;
    }
    wave_sync();
    double q[16];
#pragma unroll
    for (int c = 0; c < 16; c++)
        q[c] = rb[L * kUpRowPad + c];
    const uint64_t r = (tile << 6) + L; // this lane's row: node index at level lin+5
#pragma unroll
    for (int c = 0; c < 8; c++)
        q[c] = __dadd_rn(q[2 * c], q[2 * c + 1]);
#pragma unroll
    for (int c = 0; c < 4; c++)
        reinterpret_cast<dvec2 *>(o2)[4 * r + c] = dvec2{q[2 * c], q[2 * c + 1]};
#pragma unroll
    for (int c = 0; c < 4; c++)
        q[c] = __dadd_rn(q[2 * c], q[2 * c + 1]);
#pragma unroll
    for (int c = 0; c < 2; c++)
        reinterpret_cast<dvec2 *>(o3)[2 * r + c] = dvec2{q[2 * c], q[2 * c + 1]};
    q[0] = __dadd_rn(q[0], q[1]);
    q[1] = __dadd_rn(q[2], q[3]);
    reinterpret_cast<dvec2 *>(o4)[r] = dvec2{q[0], q[1]};
    o5[r] = __dadd_rn(q[0], q[1]);
}

// prefix_s (Fenwick over levels >= I) and tail_s = prefix_s + N_I[s].
__global__ __launch_bounds__(kThreads) void tails_kernel(const double *__restrict__ tree, const Levels lv, int I, int m,
                                                         uint64_t nseg, double *__restrict__ prefix,
                                                         double *__restrict__ tail)
{
    const uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (s >= nseg)
        return;
    double acc = 0.0;
    for (int l = m - I - 1; l >= 0; l--)
        if ((s >> l) & 1u)
            acc = __dadd_rn(acc, tree[lv.off[I + l] + ((s >> l) - 1)]);
    prefix[s] = acc;
    tail[s] = __dadd_rn(acc, tree[lv.off[I] + s]);
}

// In-place running maximum over nseg values (one workgroup; max is exact).
__global__ __launch_bounds__(kThreads) void cummax_kernel(double *__restrict__ v, uint64_t n)
{
    __shared__ double part[kThreads];
    const uint64_t per = (n + kThreads - 1) / kThreads;
    const uint64_t b = threadIdx.x * per;
    const uint64_t e = b + per < n ? b + per : n;
    double mx = -1.0;
    for (uint64_t i = b; i < e; i++)
        mx = fmax(mx, v[i]);
    part[threadIdx.x] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double run = -1.0;
        for (int t = 0; t < kThreads; t++)
        {
            const double cur = part[t];
            part[t] = run; // exclusive running max
            run = fmax(run, cur);
        }
    }
    __syncthreads();
    double run = part[threadIdx.x];
    for (uint64_t i = b; i < e; i++)
    {
        run = fmax(run, v[i]);
        v[i] = run;
    }
}

template <int N>
__device__ __forceinline__ double pick(const double (&a)[N], int idx)
{
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++)
        r = (i == idx) ? a[i] : r;
    return r;
}

// Levels below LB (<= kLBM) of the 2^LB-amplitude block holding g, as the
// block's balanced pair sums: the Fenwick part of the exclusive prefix of g's
// offset in the block (levels LB-1 .. 0, high to low) and p_g.
template <int kLBM>
__device__ __forceinline__ void block_prefix(const void *state, bool c64, uint64_t g, int LB, double &acc,
                                             double &pg)
{
    constexpr int NB = 1 << kLBM;
    const uint64_t gb = g & ~((uint64_t(1) << LB) - 1);
    const int nb = 1 << LB;
    double lv0[NB], lv1[NB / 2], lv2[NB / 4], lv3[NB / 8], lv4[NB / 16 > 0 ? NB / 16 : 1];
#pragma unroll
    for (int i = 0; i < NB; i++)
        lv0[i] = i < nb ? p_of(state, c64, gb + i) : 0.0;
#pragma unroll
    for (int i = 0; i < NB / 2; i++)
        lv1[i] = __dadd_rn(lv0[2 * i], lv0[2 * i + 1]);
#pragma unroll
    for (int i = 0; i < NB / 4; i++)
        lv2[i] = __dadd_rn(lv1[2 * i], lv1[2 * i + 1]);
#pragma unroll
    for (int i = 0; i < NB / 8; i++)
        lv3[i] = __dadd_rn(lv2[2 * i], lv2[2 * i + 1]);
    if (kLBM == 5)
    {
#pragma unroll
        for (int i = 0; i < NB / 16; i++)
            lv4[i] = __dadd_rn(lv3[2 * i], lv3[2 * i + 1]);
    }
    const int off = static_cast<int>(g - gb);
    if (kLBM == 5 && (off & 16))
        acc = __dadd_rn(acc, lv4[0]);
    if (off & 8)
        acc = __dadd_rn(acc, pick(lv3, (off >> 3) - 1));
    if (off & 4)
        acc = __dadd_rn(acc, pick(lv2, (off >> 2) - 1));
    if (off & 2)
        acc = __dadd_rn(acc, pick(lv1, (off >> 1) - 1));
    if (off & 1)
        acc = __dadd_rn(acc, pick(lv0, off - 1));
    pg = pick(lv0, off);
}

// One shot per thread.
template <bool kC64>
__global__ __launch_bounds__(kThreads) void search_kernel(const void *__restrict__ state,
                                                          const double *__restrict__ tree, const Levels lv, int I,
                                                          int LS, const double *__restrict__ cmax,
                                                          const double *__restrict__ prefix, uint64_t nseg,
                                                          const double *__restrict__ thr, uint64_t shots,
                                                          uint64_t *__restrict__ out, int *__restrict__ err)
{
    const uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (k >= shots)
        return;
    const double tau = thr[k];
    // first segment whose running-max tail exceeds tau
    uint64_t lo = 0, len = nseg;
    while (len > 0)
    {
        const uint64_t half = len >> 1;
        if (tau < cmax[lo + half])
            len = half;
        else
        {
            lo = lo + half + 1;
            len = len - half - 1;
        }
    }
    if (lo >= nseg)
    {
        *err = 1;
        out[k] = 0;
        return;
    }
    const uint64_t seg = lo;
    const uint64_t sbase = seg << I;
    const double head = prefix[seg];
    const int LB = LS < I ? LS : I; // low levels recomputed from amplitudes
    const uint64_t valid = uint64_t(1) << I;

    // libstdc++ upper_bound probe sequence over cdf[0, valid)
    uint64_t first = 0;
    len = valid;
    while (len > 0)
    {
        const uint64_t half = len >> 1;
        const uint64_t mid = first + half;
        const uint64_t g = sbase + mid;
        double acc = 0.0;
        for (int l = I - 1; l >= LB; l--)
            if ((mid >> l) & 1u)
                acc = __dadd_rn(acc, tree[lv.off[l] + ((g >> l) - 1)]);
        double pmid;
        block_prefix<kC64 ? 5 : 4>(state, kC64, g, LB, acc, pmid);
        const double cdf = __dadd_rn(__dadd_rn(head, acc), pmid);
        if (tau < cdf)
            len = half;
        else
        {
            first = mid + 1;
            len = len - half - 1;
        }
    }
    out[k] = sbase + (first == valid ? valid - 1 : first);
}

// Lowest stored tree level: below it the search recomputes from amplitudes
// (256 bytes of state per block in both precisions).
inline unsigned low_level(int prec, unsigned I)
{
    return std::min(prec == C64 ? 5u : 4u, I);
}

int build_tree(Ctx &ctx, const void *state, int prec, unsigned m, unsigned LS, Levels &lv, double **tree_out)
{
    uint64_t total = 0;
    for (unsigned l = 0; l < kMaxLevels; l++)
        lv.off[l] = 0;
    for (unsigned l = LS; l <= m; l++)
    {
        lv.off[l] = total;
        total += uint64_t(1) << (m - l);
    }
    ZXA_TRY(ctx.scratch_b.ensure(total * sizeof(double)));
    double *tree = static_cast<double *>(ctx.scratch_b.ptr);
    // level 0 -> up to kSpanBits, then from level kSpanBits, ...
    int lin = 0;
    bool first = true;
    ZXA_TRY(ctx.tic());
    const int gbits = prec == C64 ? TreeGeo<true>::kGroupBits : TreeGeo<false>::kGroupBits;
    const unsigned amp_level = prec == C64 ? 5u : 4u;
    if (LS == amp_level && int(m) >= gbits)
    {
        // amplitudes -> level 4 (c128) or 5 (c64) in one read of the state
        const unsigned groups = static_cast<unsigned>(uint64_t(1) << (int(m) - gbits));
        double *tl = tree + lv.off[amp_level];
        if (prec == C64)
            hipLaunchKernelGGL(tree_amp_kernel<true>, dim3(groups), dim3(kThreads), 0, ctx.stream, state, tl);
        else
            hipLaunchKernelGGL(tree_amp_kernel<false>, dim3(groups), dim3(kThreads), 0, ctx.stream, state, tl);
        ZXA_HIP(hipGetLastError());
        first = false;
        lin = int(amp_level);
    }
    // stored levels -> five more levels per pass while a pass fills its grid
    while (!first && int(m) - lin >= 13 + 5)
    {
        const unsigned groups = static_cast<unsigned>(uint64_t(1) << (int(m) - lin - 13));
        hipLaunchKernelGGL(tree_up_kernel, dim3(groups), dim3(kThreads), 0, ctx.stream, tree + lv.off[lin],
                           tree + lv.off[lin + 1], tree + lv.off[lin + 2], tree + lv.off[lin + 3],
                           tree + lv.off[lin + 4], tree + lv.off[lin + 5]);
        ZXA_HIP(hipGetLastError());
        lin += 5;
    }
    while (first || lin < int(m))
    {
        const int span_bits = std::min<int>(kSpanBits, int(m) - lin);
        const uint64_t groups = uint64_t(1) << (int(m) - lin - span_bits);
        const int kind = first ? (prec == C64 ? 1 : 0) : 2;
        const void *src = first ? state : static_cast<const void *>(tree);
        // a non-first pass reads level lin, which must be stored
        if (!first && unsigned(lin) < LS)
        {
            set_error("sampler tree: unstored intermediate level");
            return EINVAL_;
        }
        hipLaunchKernelGGL(tree_kernel, dim3(static_cast<unsigned>(groups)), dim3(kThreads), 0, ctx.stream, src, kind,
                           lin, span_bits, int(LS), int(m), tree, lv);
        ZXA_HIP(hipGetLastError());
        first = false;
        lin += span_bits;
        if (span_bits == 0)
            break;
    }
    // algorithmic bytes: one read of the state and one write of the stored
    // levels (the passes above the first read stored levels back; not counted)
    ZXA_TRY(ctx.toc(F_TREE, (uint64_t(1) << m) * elem_bytes(prec) + total * sizeof(double)));
    *tree_out = tree;
    return OK;
}

} // namespace

int prob_total(Ctx &ctx, const void *state, int prec, unsigned m, double *out)
{
    Levels lv;
    double *tree = nullptr;
    const unsigned LS = low_level(prec, m);
    ZXA_TRY(build_tree(ctx, state, prec, m, LS, lv, &tree));
    ZXA_HIP(hipMemcpyAsync(out, tree + lv.off[m], sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    ZXA_HIP(hipStreamSynchronize(ctx.stream));
    return OK;
}

int sample_launch(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits)
{
    if (m > 40)
    {
        set_error("sampler: state wider than 40 bits");
        return EINVAL_;
    }
    ctx.sample_ready = false;
    const unsigned I = std::min(m, seg_bits);
    const unsigned LS = low_level(prec, I);
    Levels &lv = ctx.sample_levels_as<Levels>();
    double *tree = nullptr;
    ZXA_TRY(ctx.sample_total_pin.ensure(sizeof(double)));
    ZXA_TRY(build_tree(ctx, state, prec, m, LS, lv, &tree));
    const uint64_t nseg = uint64_t(1) << (m - I);
    ZXA_TRY(ctx.scratch_d.ensure(2 * nseg * sizeof(double)));
    double *prefix = static_cast<double *>(ctx.scratch_d.ptr);
    double *cmax = prefix + nseg;
    hipLaunchKernelGGL(tails_kernel, dim3(static_cast<unsigned>((nseg + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       ctx.stream, tree, lv, int(I), int(m), nseg, prefix, cmax);
    ZXA_HIP(hipGetLastError());
    hipLaunchKernelGGL(cummax_kernel, dim3(1), dim3(kThreads), 0, ctx.stream, cmax, nseg);
    ZXA_HIP(hipGetLastError());
    ZXA_HIP(hipMemcpyAsync(ctx.sample_total_pin.ptr, tree + lv.off[m], sizeof(double), hipMemcpyDeviceToHost,
                           ctx.stream));
    ctx.sample_m = m;
    ctx.sample_I = I;
    ctx.sample_ready = true;
    return OK;
}

int sample_total(Ctx &ctx, double *total)
{
    if (!ctx.sample_ready)
    {
        set_error("sampler: total before launch");
        return ESTATE;
    }
    ZXA_HIP(hipStreamSynchronize(ctx.stream));
    *total = *static_cast<const double *>(ctx.sample_total_pin.ptr);
    return OK;
}

int sample_prepare(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits, double *total)
{
    ZXA_TRY(sample_launch(ctx, state, prec, m, seg_bits));
    return sample_total(ctx, total);
}

// Real address -> virtual result words (pack_virtual_results_host,
// src/zxh/src/measure/measure_result_pack.h:24-44; the reference's GPU path
// does the same in map_real_to_virtual_words_kernel): virtual bit q of shot i
// = parity(real[i] & rows[q]) ^ b[q], packed into wc 64-bit words per shot.
__global__ __launch_bounds__(kThreads) void pack_kernel(const uint64_t *__restrict__ real, uint64_t shots,
                                                        const uint64_t *__restrict__ rows,
                                                        const uint8_t *__restrict__ b, int nq, int wc,
                                                        uint64_t *__restrict__ out)
{
    const uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= shots)
        return;
    const uint64_t r = real[i];
    for (int w = 0; w < wc; w++)
    {
        uint64_t word = 0;
        const int hi = min(64, nq - 64 * w);
        for (int k = 0; k < hi; k++)
        {
            const int q = 64 * w + k;
            word |= uint64_t((__popcll(r & rows[q]) & 1) ^ b[q]) << k;
        }
        out[i * uint64_t(wc) + w] = word;
    }
}

int sample_search(Ctx &ctx, const void *state, int prec, const double *thresholds, size_t shots, uint64_t *out_real,
                  const PackSpec *pack)
{
    if (shots == 0)
        return OK;
    if (!ctx.sample_ready)
    {
        set_error("sampler: search before prepare");
        return ESTATE;
    }
    const unsigned m = ctx.sample_m, I = ctx.sample_I;
    const unsigned LS = low_level(prec, I);
    const Levels &lv = ctx.sample_levels_as<Levels>();
    const double *tree = static_cast<const double *>(ctx.scratch_b.ptr);
    const uint64_t nseg = uint64_t(1) << (m - I);
    const double *prefix = static_cast<const double *>(ctx.scratch_d.ptr);
    const double *cmax = prefix + nseg;
    const size_t wc = pack ? (size_t(pack->nq) + 63) / 64 : 0;
    const size_t nq = pack ? pack->nq : 0;
    // thresholds | real results | packed words | rows | error flag | b
    const size_t need = shots * (sizeof(double) + sizeof(uint64_t) * (1 + wc)) + nq * sizeof(uint64_t) + 64 + nq;
    ZXA_TRY(ctx.scratch_c.ensure(need));
    double *dthr = static_cast<double *>(ctx.scratch_c.ptr);
    uint64_t *dout = reinterpret_cast<uint64_t *>(dthr + shots);
    uint64_t *dpacked = dout + shots;
    uint64_t *drows = dpacked + shots * wc;
    int *derr = reinterpret_cast<int *>(drows + nq);
    uint8_t *db = reinterpret_cast<uint8_t *>(derr + 16);
    ZXA_HIP(hipMemcpyAsync(dthr, thresholds, shots * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
    ZXA_HIP(hipMemsetAsync(derr, 0, sizeof(int), ctx.stream));
    const unsigned grid = static_cast<unsigned>((shots + kThreads - 1) / kThreads);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(search_kernel<true>, dim3(grid), dim3(kThreads), 0, ctx.stream, state, tree, lv, int(I),
                           int(LS), cmax, prefix, nseg, dthr, uint64_t(shots), dout, derr);
    else
        hipLaunchKernelGGL(search_kernel<false>, dim3(grid), dim3(kThreads), 0, ctx.stream, state, tree, lv, int(I),
                           int(LS), cmax, prefix, nseg, dthr, uint64_t(shots), dout, derr);
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_SAMPLE, 0));
    if (pack)
    {
        ZXA_HIP(hipMemcpyAsync(drows, pack->rows, nq * sizeof(uint64_t), hipMemcpyHostToDevice, ctx.stream));
        ZXA_HIP(hipMemcpyAsync(db, pack->b, nq, hipMemcpyHostToDevice, ctx.stream));
        hipLaunchKernelGGL(pack_kernel, dim3(grid), dim3(kThreads), 0, ctx.stream, dout, uint64_t(shots), drows, db,
                           int(nq), int(wc), dpacked);
        ZXA_HIP(hipGetLastError());
        ZXA_HIP(hipMemcpyAsync(pack->out, dpacked, shots * wc * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               ctx.stream));
    }
    int herr = 0;
    ZXA_HIP(hipMemcpyAsync(out_real, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx.stream));
    ZXA_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
    ZXA_HIP(hipStreamSynchronize(ctx.stream));
    if (herr)
    {
        set_error("measure failed to resolve all local samples");
        return EUNRESOLVED;
    }
    return OK;
}

// measure (measure_common.cpp:73-92) on a caller's state and affine map:
// the tree builds while the host draws the threshold stream's uniforms and
// factors, the search resolves the thresholds, and the device packs the
// virtual result words.
int measure_affine(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits, unsigned nq,
                   const uint64_t *rows, const uint8_t *b, bool use_seed, uint64_t seed, size_t shots,
                   uint64_t *packed, uint64_t *out_real)
{
    if (shots == 0)
        return OK;
    ZXA_TRY(sample_launch(ctx, state, prec, m, std::min(m, seg_bits)));
    const uint64_t base = use_seed ? seed : uint64_t(std::random_device{}());
    ZXA_TRY(ctx.sample_thr_pin.ensure(shots * sizeof(double)));
    double *thr = static_cast<double *>(ctx.sample_thr_pin.ptr);
    threshold_draws(mix_seed(base, 0), shots, thr);
    threshold_factors(thr, shots);
    double total = 0.0;
    ZXA_TRY(sample_total(ctx, &total));
    if (!(total > 0.0))
    {
        set_error("state has zero total probability");
        return EZEROPROB;
    }
    threshold_apply(thr, shots, 0.0, total);
    std::vector<uint64_t> realv;
    if (!out_real)
    {
        realv.resize(shots);
        out_real = realv.data();
    }
    PackSpec ps;
    ps.nq = nq;
    ps.rows = rows;
    ps.b = b;
    ps.out = packed;
    return sample_search(ctx, state, prec, thr, shots, out_real, &ps);
}

int sample_apply(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits, const double *thresholds,
                 size_t shots, uint64_t *out_real)
{
    if (shots == 0)
        return OK;
    double total = 0.0;
    ZXA_TRY(sample_prepare(ctx, state, prec, m, seg_bits, &total));
    return sample_search(ctx, state, prec, thresholds, shots, out_real);
}

} // namespace zxa
