// histogram_even for gfx950, and the zero-fill kernel the other entry
// points use to clear scratch.
//
// Contract of the reference's detail::histogram_even
// (lib/src/Compute/Algorithm.cpp:240-273, :569-615): zero num_bins counters,
// count x with lower <= x < upper into bin (uint)((x - lower) * scale),
// scale = (ACC)bins / (upper - lower), clamped to bins - 1, where ACC is
// float unless the samples are double and lower/upper are first rounded to
// ACC. Each workgroup privatises the bins in LDS when they fit (64 KiB of
// counters), reads 16 bytes per lane per load, and adds its non-zero bins to
// the global histogram once.
#include "context.hpp"
#include "device_common.hpp"
#include "ops.hpp"
#include "scan_kernels.hpp"

#include <algorithm>

namespace clc {
namespace {

using namespace dev;

__global__ __launch_bounds__(256) void zero_kernel(uint4* p, size_t n16,
                                                   unsigned char* tail,
                                                   size_t ntail)
{
    const size_t stride = size_t(gridDim.x) * 256;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

template <typename T, typename ACC, bool LOCAL, int BLOCK>
__global__ __launch_bounds__(BLOCK) void histogram_kernel(
    const T* __restrict__ in, size_t n, uint32_t* __restrict__ hist,
    uint32_t bins, ACC lower, ACC upper, ACC scale, int vec_ok)
{
    extern __shared__ uint32_t s_bins[];
    if constexpr (LOCAL)
    {
        for (uint32_t b = threadIdx.x; b < bins; b += BLOCK) s_bins[b] = 0;
        __syncthreads();
    }
    auto count = [&](T v) {
        const ACC x = static_cast<ACC>(v);
        if (x >= lower && x < upper)
        {
            uint32_t b = static_cast<uint32_t>((x - lower) * scale);
            if (b >= bins) b = bins - 1;
            if constexpr (LOCAL) atomicAdd(&s_bins[b], 1u);
            else atomicAdd(&hist[b], 1u);
        }
    };
    const size_t stride = size_t(gridDim.x) * BLOCK;
    size_t start = 0;
    if (vec_ok)
    {
        constexpr int VEC = 16 / sizeof(T);
        const size_t nvec = n / VEC;
        const uint4* v = reinterpret_cast<const uint4*>(in);
        for (size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nvec; i += stride)
        {
            const uint4 q = load_stream(v + i);
            T e[VEC];
            __builtin_memcpy(e, &q, 16);
#pragma unroll
            for (int j = 0; j < VEC; ++j) count(e[j]);
        }
        start = nvec * VEC;
    }
    for (size_t i = start + size_t(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += stride)
        count(in[i]);
    if constexpr (LOCAL)
    {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < bins; b += BLOCK)
            if (s_bins[b] != 0) atomicAdd(&hist[b], s_bins[b]);
    }
}

template <typename T, typename ACC>
void histogram_typed(context& ctx, const void* in, size_t n, uint32_t* hist,
                     uint32_t bins, double lower, double upper)
{
    constexpr int BLOCK = 256;
    const ACC lo = static_cast<ACC>(lower), up = static_cast<ACC>(upper);
    // Computed in ACC on the host: IEEE division, correctly rounded.
    const ACC scale = static_cast<ACC>(bins) / (up - lo);
    const size_t lds = size_t(bins) * 4;
    const bool local = lds + 1024 <= ctx.facts().lds_per_block;
    const size_t per_block = size_t(BLOCK) * 64;
    const uint32_t groups = static_cast<uint32_t>(std::max<size_t>(
        1, std::min<size_t>((n + per_block - 1) / per_block,
                            size_t(ctx.facts().cu_count) * 8)));
    const int vec_ok = reinterpret_cast<uintptr_t>(in) % 16 == 0 ? 1 : 0;
    const T* src = static_cast<const T*>(in);
    timed_launch tl(ctx, "histogram_even");
    if (local)
        hipLaunchKernelGGL((histogram_kernel<T, ACC, true, BLOCK>), dim3(groups),
                           dim3(BLOCK), lds, ctx.stream(), src, n, hist, bins,
                           lo, up, scale, vec_ok);
    else
        hipLaunchKernelGGL((histogram_kernel<T, ACC, false, BLOCK>), dim3(groups),
                           dim3(BLOCK), 0, ctx.stream(), src, n, hist, bins, lo,
                           up, scale, vec_ok);
    CLC_HIP(hipGetLastError());
}

} // namespace

namespace detail {

void zero_fill(hipStream_t stream, void* p, std::size_t bytes)
{
    if (bytes == 0) return;
    const size_t n16 = bytes / 16;
    const size_t ntail = bytes % 16;
    const uint32_t grid = static_cast<uint32_t>(
        std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 4096)));
    hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, stream,
                       static_cast<uint4*>(p), n16,
                       static_cast<unsigned char*>(p) + n16 * 16, ntail);
    CLC_HIP(hipGetLastError());
}

void histogram_even(context& ctx, const void* in, std::size_t n, int type,
                    uint32_t* histogram, std::size_t num_bins, double lower,
                    double upper)
{
    if (num_bins == 0) return;
    if (!(lower < upper))
        throw std::invalid_argument("histogram_even: lower must be below upper");
    check_size(n);
    if (num_bins > 0xFFFFFFFFu)
        throw std::invalid_argument("histogram_even: too many bins");
    if (histogram == nullptr)
        throw std::invalid_argument("histogram_even: null histogram");
    type_size(type); // validates the type code
    ctx.activate();
    zero_fill(ctx.stream(), histogram, num_bins * 4);
    if (n == 0) return;
    if (in == nullptr) throw std::invalid_argument("histogram_even: null input");
    const uint32_t bins = static_cast<uint32_t>(num_bins);
    switch (type)
    {
        case TY_I8: histogram_typed<int8_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_U8: histogram_typed<uint8_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_I16: histogram_typed<int16_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_U16: histogram_typed<uint16_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_I32: histogram_typed<int32_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_U32: histogram_typed<uint32_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_I64: histogram_typed<int64_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_U64: histogram_typed<uint64_t, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_F32: histogram_typed<float, float>(ctx, in, n, histogram, bins, lower, upper); return;
        case TY_F64: histogram_typed<double, double>(ctx, in, n, histogram, bins, lower, upper); return;
    }
}

} // namespace detail
} // namespace clc
