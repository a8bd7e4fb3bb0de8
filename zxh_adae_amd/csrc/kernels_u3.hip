// kernels_u3.hip -- general pair sweeps, permutations and the CP fallback.
//
// Single-bit U3 gates whose selector contains their target bit run as tile
// passes (kernels_tile.hip).  Gates whose pair offset delta has several bits
// (CX transport in the ZXH affine mode), and states narrower than a tile, use
// the pair kernel here: one sweep per gate, pairs enumerated by inserting a
// zero at the lowest delta bit (u3_block_batch_kernel_cuda,
// src/zxh/src/kernel/kernel_cuda.cu:390-423, per descriptor).
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

// The U3 that expands a qubit: its pair offset is the new top bit, whose half
// of the state is still zero, so each pair reads one amplitude a and writes
// (m00 a, m10 a) (a on the even side) or (m11 a, m01 a) (a on the odd side):
// one read and two writes per pair instead of two and two.
template <bool kC64>
__global__ __launch_bounds__(kThreads) void u3_expand_kernel(void *__restrict__ state, uint64_t half,
                                                             uint64_t packed, double u00x, double u00y, double u01x,
                                                             double u01y, double u10x, double u10y, double u11x,
                                                             double u11y)
{
    using T = typename Elem<kC64>::T;
    using R = typename Elem<kC64>::R;
    T *p = static_cast<T *>(state);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < half; k += stride)
    {
        const bool odd = sel_eval(packed, k);
        const T a = p[k];
        // (this amplitude's new value, its partner's) = (m00, m10) a or (m11, m01) a
        const R sx = R(odd ? u11x : u00x), sy = R(odd ? u11y : u00y);
        const R tx = R(odd ? u01x : u10x), ty = R(odd ? u01y : u10y);
        T na, nb;
        na.x = sx * a.x - sy * a.y;
        na.y = sx * a.y + sy * a.x;
        nb.x = tx * a.x - ty * a.y;
        nb.y = tx * a.y + ty * a.x;
        p[k] = na;
        p[k + half] = nb;
    }
}

// A run of J (2..4) expansions at once.  Gate i expands address bit
// base + i on the state of base + i bits; its selector may use every lower
// bit.  One thread per old amplitude a[k] builds the 2^J new amplitudes
// k + y 2^base in registers, doubling per gate as the expansions one at a
// time do: with odd = S_i(k + y 2^base) for the y built so far, the
// amplitude v becomes m00 v (bit base + i zero) and m10 v (one), or m11 v
// and m01 v when odd.  Thread k alone reads a[k], so writing in place is safe.
struct ExpGate
{
    uint64_t packed;
    double u[8];
};

template <bool kC64, int J>
__global__ __launch_bounds__(kThreads) void u3_expand_group_kernel(void *__restrict__ state, unsigned base,
                                                                   uint64_t old, const ExpGate *__restrict__ gates)
{
    using T = typename Elem<kC64>::T;
    using R = typename Elem<kC64>::R;
    T *p = static_cast<T *>(state);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; k < old; k += stride)
    {
        R vx[1 << J], vy[1 << J];
        const T a = p[k];
        vx[0] = a.x;
        vy[0] = a.y;
#pragma unroll
        for (int i = 0; i < J; i++)
        {
#pragma unroll
            for (int y = 0; y < (1 << i); y++)
            {
                const bool odd = sel_eval(gates[i].packed, k | (uint64_t(y) << base));
                const int a0 = odd ? 6 : 0, a1 = odd ? 2 : 4;
                const R s0x = R(gates[i].u[a0]), s0y = R(gates[i].u[a0 + 1]);
                const R s1x = R(gates[i].u[a1]), s1y = R(gates[i].u[a1 + 1]);
                const R x = vx[y], w = vy[y];
                vx[y | (1 << i)] = s1x * x - s1y * w;
                vy[y | (1 << i)] = s1x * w + s1y * x;
                vx[y] = s0x * x - s0y * w;
                vy[y] = s0x * w + s0y * x;
            }
        }
#pragma unroll
        for (int y = 0; y < (1 << J); y++)
        {
            T out;
            out.x = vx[y];
            out.y = vy[y];
            p[k + (uint64_t(y) << base)] = out;
        }
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

// Zero fill of 2^(nb + 9) 16-byte units: workgroup w writes chunk rot(w) of
// 512 units (two per thread) with nontemporal stores, the chunk index rotated
// right by two bits so that consecutive workgroups write a quarter of the
// buffer apart (on a 16 GiB state: 2.57 ms, against 2.81 ms for a linear
// order or a runtime memset).
typedef double zvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kThreads) void zero_kernel(zvec2 *p, unsigned nb)
{
    const uint64_t c = blockIdx.x;
    const uint64_t mask = (uint64_t(1) << nb) - 1;
    const uint64_t ch = nb >= 2 ? (((c >> 2) | (c << (nb - 2))) & mask) : c;
    zvec2 *q = p + ch * 512 + threadIdx.x;
    const zvec2 z = {0.0, 0.0};
    __builtin_nontemporal_store(z, q);
    __builtin_nontemporal_store(z, q + 256);
}

// Zero fill of n doubles (any n), grid-stride.
__global__ __launch_bounds__(kThreads) void zero_any_kernel(double *p, uint64_t n)
{
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride)
        p[i] = 0.0;
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

int u3_expand_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate &g)
{
    if (n == 0 || g.delta != (uint64_t(1) << (n - 1)))
    {
        set_error("u3 expand: the pair offset is not the state's top bit");
        return EINVAL_;
    }
    const uint64_t half = uint64_t(1) << (n - 1);
    const uint64_t bytes = 3ull * half * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(u3_expand_kernel<true>, dim3(grid_for(half)), dim3(kThreads), 0, ctx.stream, state, half,
                           g.packed_sel, g.u[0], g.u[1], g.u[2], g.u[3], g.u[4], g.u[5], g.u[6], g.u[7]);
    else
        hipLaunchKernelGGL(u3_expand_kernel<false>, dim3(grid_for(half)), dim3(kThreads), 0, ctx.stream, state, half,
                           g.packed_sel, g.u[0], g.u[1], g.u[2], g.u[3], g.u[4], g.u[5], g.u[6], g.u[7]);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_U3PAIR, bytes);
}

int u3_expand_group_apply(Ctx &ctx, void *state, int prec, unsigned base, const U3Gate *gates, size_t count)
{
    if (count == 0)
        return OK;
    if (count == 1)
        return u3_expand_apply(ctx, state, prec, base + 1, gates[0]);
    if (count > 4 || base + count > 40)
    {
        set_error("u3 expand group: at most four expansions, within 40 bits");
        return EINVAL_;
    }
    std::vector<ExpGate> tab(count);
    for (size_t i = 0; i < count; i++)
    {
        const unsigned b = base + unsigned(i);
        const uint64_t alpha = gates[i].packed_sel & ((uint64_t(1) << 63) - 1);
        if (gates[i].delta != (uint64_t(1) << b) || (alpha >> (b + 1)) != 0)
        {
            set_error("u3 expand group: gate i must expand bit base + i and select below it");
            return EINVAL_;
        }
        tab[i].packed = gates[i].packed_sel;
        std::memcpy(tab[i].u, gates[i].u, sizeof(tab[i].u));
    }
    void *dt = nullptr;
    ZXA_TRY(ctx.ring.stage(tab.data(), tab.size() * sizeof(ExpGate), ctx.stream, &dt));
    const ExpGate *dg = static_cast<const ExpGate *>(dt);
    const uint64_t old = uint64_t(1) << base;
    // bytes: the old amplitudes read, the final ones written
    const uint64_t bytes = (old + (old << count)) * elem_bytes(prec);
    const dim3 grid(grid_for(old)), block(kThreads);
    ZXA_TRY(ctx.tic());
    if (prec == C64)
    {
        if (count == 2)
            hipLaunchKernelGGL((u3_expand_group_kernel<true, 2>), grid, block, 0, ctx.stream, state, base, old, dg);
        else if (count == 3)
            hipLaunchKernelGGL((u3_expand_group_kernel<true, 3>), grid, block, 0, ctx.stream, state, base, old, dg);
        else
            hipLaunchKernelGGL((u3_expand_group_kernel<true, 4>), grid, block, 0, ctx.stream, state, base, old, dg);
    }
    else
    {
        if (count == 2)
            hipLaunchKernelGGL((u3_expand_group_kernel<false, 2>), grid, block, 0, ctx.stream, state, base, old, dg);
        else if (count == 3)
            hipLaunchKernelGGL((u3_expand_group_kernel<false, 3>), grid, block, 0, ctx.stream, state, base, old, dg);
        else
            hipLaunchKernelGGL((u3_expand_group_kernel<false, 4>), grid, block, 0, ctx.stream, state, base, old, dg);
    }
    ZXA_HIP(hipGetLastError());
    ZXA_TRY(ctx.toc(F_U3PAIR, bytes));
    return ctx.ring.done(ctx.stream);
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
    const uint64_t bytes = uint64_t(alloc_elems) * elem_bytes(prec);
    const uint64_t units = bytes / 16;
    if (bytes % 16 == 0 && units >= 512 && (units & (units - 1)) == 0 && units / 512 <= uint64_t(INT_MAX))
    {
        const unsigned nb = unsigned(__builtin_ctzll(units)) - 9;
        hipLaunchKernelGGL(zero_kernel, dim3(unsigned(units >> 9)), dim3(kThreads), 0, ctx.stream,
                           static_cast<zvec2 *>(state), nb);
    }
    else
    {
        const uint64_t nd = bytes / 8; // whole doubles: both precisions are multiples of 8 bytes
        const unsigned grid = unsigned(std::min<uint64_t>((nd + kThreads - 1) / kThreads, 65536));
        hipLaunchKernelGGL(zero_any_kernel, dim3(std::max(1u, grid)), dim3(kThreads), 0, ctx.stream,
                           static_cast<double *>(state), nd);
    }
    if (prec == C64)
        hipLaunchKernelGGL(set_one_kernel<true>, dim3(1), dim3(1), 0, ctx.stream, state);
    else
        hipLaunchKernelGGL(set_one_kernel<false>, dim3(1), dim3(1), 0, ctx.stream, state);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_RESET, uint64_t(alloc_elems) * elem_bytes(prec));
}

} // namespace zxa
