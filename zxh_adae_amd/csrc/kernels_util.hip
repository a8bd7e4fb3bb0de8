// kernels_util.hip -- streaming helpers: a device copy (the bandwidth and
// counter calibration reference) and an in-place complex scale (renormalise /
// apply a global phase).
#include "dev.h"

#include <algorithm>

namespace zxa
{
namespace
{

constexpr int kThreads = 256;
constexpr int kUnroll = 8;

// 16-byte vectors, kUnroll independent loads in flight per thread.
__global__ __launch_bounds__(kThreads) void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                        uint64_t n16)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads * kUnroll;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kThreads * kUnroll + threadIdx.x; base < n16;
         base += stride)
    {
        float4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
        {
            const uint64_t i = base + uint64_t(u) * kThreads;
            if (i < n16)
                v[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
        {
            const uint64_t i = base + uint64_t(u) * kThreads;
            if (i < n16)
                dst[i] = v[u];
        }
    }
}

// One 16-byte load per thread across as many blocks as the block needs: the
// source is host memory read over the bus, so every load is issued at once
// rather than a few per thread in sequence.
__global__ __launch_bounds__(kThreads) void ring_copy_kernel(const float4 *__restrict__ src,
                                                             float4 *__restrict__ dst, uint64_t n16)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n16)
        dst[i] = src[i];
}

template <bool kC64>
__global__ __launch_bounds__(kThreads) void scale_kernel(void *__restrict__ state, uint64_t n16, double re, double im)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads * kUnroll;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kThreads * kUnroll + threadIdx.x; base < n16;
         base += stride)
    {
        if constexpr (!kC64)
        {
            double2 *p = static_cast<double2 *>(state);
            double2 v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
            {
                const uint64_t i = base + uint64_t(u) * kThreads;
                if (i < n16)
                    v[u] = p[i];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
            {
                const uint64_t i = base + uint64_t(u) * kThreads;
                if (i < n16)
                    p[i] = make_double2(v[u].x * re - v[u].y * im, v[u].x * im + v[u].y * re);
            }
        }
        else
        {
            float4 *p = static_cast<float4 *>(state);
            const float r = static_cast<float>(re), m = static_cast<float>(im);
            float4 v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
            {
                const uint64_t i = base + uint64_t(u) * kThreads;
                if (i < n16)
                    v[u] = p[i];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
            {
                const uint64_t i = base + uint64_t(u) * kThreads;
                if (i < n16)
                    p[i] = make_float4(v[u].x * r - v[u].y * m, v[u].x * m + v[u].y * r, v[u].z * r - v[u].w * m,
                                       v[u].z * m + v[u].w * r);
            }
        }
    }
}

// Zero fill of n16 16-byte units (graph-capturable, unlike a memset node).
__global__ __launch_bounds__(kThreads) void zero_kernel(float4 *__restrict__ p, uint64_t n16)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n16; i += stride)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Amplitude 0 of each listed state = 1.
template <bool kC64>
__global__ void set_ones_kernel(const uint64_t *__restrict__ states, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    if constexpr (kC64)
        reinterpret_cast<float2 *>(states[i])[0] = make_float2(1.f, 0.f);
    else
        reinterpret_cast<double2 *>(states[i])[0] = make_double2(1.0, 0.0);
}

// out[i] = amplitude idx[i] (indices checked on the host).
template <bool kC64>
__global__ __launch_bounds__(kThreads) void gather_kernel(const void *__restrict__ state,
                                                          const uint64_t *__restrict__ idx, uint64_t count,
                                                          void *__restrict__ out)
{
    const uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= count)
        return;
    if constexpr (kC64)
        static_cast<float2 *>(out)[i] = static_cast<const float2 *>(state)[idx[i]];
    else
        static_cast<double2 *>(out)[i] = static_cast<const double2 *>(state)[idx[i]];
}

template <typename K>
unsigned stream_grid(Ctx &ctx, K kernel, uint64_t n16)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    const uint64_t cap = uint64_t(per_cu) * uint64_t(ctx.num_cus > 0 ? ctx.num_cus : 256);
    const uint64_t need = (n16 + uint64_t(kThreads) * kUnroll - 1) / (uint64_t(kThreads) * kUnroll);
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min(cap, need)));
}

} // namespace

int copy_apply(Ctx &ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes % 16 != 0)
    {
        set_error("copy: size must be a multiple of 16 bytes");
        return EINVAL_;
    }
    const uint64_t n16 = bytes / 16;
    ZXA_TRY(ctx.tic());
    hipLaunchKernelGGL(copy_kernel, dim3(stream_grid(ctx, copy_kernel, n16)), dim3(kThreads), 0, ctx.stream,
                       static_cast<const float4 *>(src), static_cast<float4 *>(dst), n16);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_COPY, 2ull * bytes);
}

int ring_copy(hipStream_t stream, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0)
        return OK;
    const uint64_t n16 = (bytes + 15) / 16;
    const unsigned grid = unsigned((n16 + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(ring_copy_kernel, dim3(grid), dim3(kThreads), 0, stream, static_cast<const float4 *>(src),
                       static_cast<float4 *>(dst), n16);
    return hip_check(hipGetLastError(), "ring copy launch");
}

int scale_apply(Ctx &ctx, void *state, int prec, unsigned n, double re, double im)
{
    const uint64_t bytes = (uint64_t(1) << n) * elem_bytes(prec);
    if (bytes < 16)
    {
        set_error("scale: state smaller than 16 bytes");
        return EINVAL_;
    }
    const uint64_t n16 = bytes / 16;
    ZXA_TRY(ctx.tic());
    if (prec == C64)
        hipLaunchKernelGGL(scale_kernel<true>, dim3(stream_grid(ctx, scale_kernel<true>, n16)), dim3(kThreads), 0,
                           ctx.stream, state, n16, re, im);
    else
        hipLaunchKernelGGL(scale_kernel<false>, dim3(stream_grid(ctx, scale_kernel<false>, n16)), dim3(kThreads), 0,
                           ctx.stream, state, n16, re, im);
    ZXA_HIP(hipGetLastError());
    return ctx.toc(F_SCALE, 2ull * bytes);
}

int zero_fill(Ctx &ctx, void *p, size_t bytes)
{
    if (bytes % 16)
    {
        set_error("zero_fill: size not a multiple of 16 bytes");
        return EINVAL_;
    }
    const uint64_t n16 = bytes / 16;
    const unsigned grid = unsigned(std::max<uint64_t>(1, std::min<uint64_t>((n16 + kThreads - 1) / kThreads, 1u << 16)));
    hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(kThreads), 0, ctx.stream, static_cast<float4 *>(p), n16);
    return hip_check(hipGetLastError(), "zero_kernel launch");
}

int gather(Ctx &ctx, const void *state, int prec, unsigned n, const uint64_t *idx, size_t count, void *out)
{
    for (size_t i = 0; i < count; i++)
        if (idx[i] >> n)
        {
            set_error("gather: index outside the state");
            return EINVAL_;
        }
    if (count == 0)
        return OK;
    const size_t esz = prec == C64 ? 8 : 16;
    void *didx = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&didx, count * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipMalloc(&dout, count * esz);
    if (e == hipSuccess)
        e = hipMemcpyAsync(didx, idx, count * sizeof(uint64_t), hipMemcpyHostToDevice, ctx.stream);
    if (e == hipSuccess)
    {
        const unsigned grid = unsigned((count + kThreads - 1) / kThreads);
        if (prec == C64)
            hipLaunchKernelGGL(gather_kernel<true>, dim3(grid), dim3(kThreads), 0, ctx.stream, state,
                               static_cast<const uint64_t *>(didx), uint64_t(count), dout);
        else
            hipLaunchKernelGGL(gather_kernel<false>, dim3(grid), dim3(kThreads), 0, ctx.stream, state,
                               static_cast<const uint64_t *>(didx), uint64_t(count), dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, count * esz, hipMemcpyDeviceToHost, ctx.stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx.stream);
    if (didx)
        (void)hipFree(didx);
    if (dout)
        (void)hipFree(dout);
    return hip_check(e, "gather");
}

int set_ones(Ctx &ctx, const uint64_t *dstates, int count, int prec)
{
    if (count <= 0)
        return OK;
    const unsigned grid = unsigned((count + kThreads - 1) / kThreads);
    if (prec == C64)
        hipLaunchKernelGGL(set_ones_kernel<true>, dim3(grid), dim3(kThreads), 0, ctx.stream, dstates, count);
    else
        hipLaunchKernelGGL(set_ones_kernel<false>, dim3(grid), dim3(kThreads), 0, ctx.stream, dstates, count);
    return hip_check(hipGetLastError(), "set_ones_kernel launch");
}

} // namespace zxa
