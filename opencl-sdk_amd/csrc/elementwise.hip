// fill, sequence, transform and transform2 for gfx950: the elementwise
// helpers of the reference's Compute library (detail::fill / sequence /
// transform / transform2, lib/src/Compute/Algorithm.cpp:331-398; C entry
// points CApi.cpp:301-369 area, compute_c.h:142-155) that its copy_if and
// histogram callers lean on.
//
// fill and sequence take no expression and are compiled ahead of time:
// fill by element width (it copies bits), sequence per element type
// (out[i] = start + (T)i * step, integer arithmetic wrapping as in OpenCL
// C, floating point as two correctly rounded operations, not a fused
// multiply-add). transform / transform2 paste the caller's expression into
// the kernel bodies of elementwise_kernels.hpp and compile them at run time
// through the context's program cache, as the reference compiles EXPR.
#include "context.hpp"
#include "elementwise_kernels.hpp"
#include "ops.hpp"
#include "rtc.hpp"

#include <cstring>

namespace clc {
namespace {

using namespace dev;

template <typename U>
__global__ __launch_bounds__(EW_BLOCK) void fill_kernel(U* __restrict__ out, size_t n,
                                                        U value)
{
    const size_t base = size_t(blockIdx.x) * EW_BLOCK * EW_ITEMS + threadIdx.x;
#pragma unroll
    for (int k = 0; k < EW_ITEMS; ++k)
    {
        const size_t i = base + size_t(k) * EW_BLOCK;
        if (i < n) out[i] = value;
    }
}

template <typename T> struct wrap_of { using type = T; };
template <> struct wrap_of<int8_t> { using type = uint32_t; };
template <> struct wrap_of<uint8_t> { using type = uint32_t; };
template <> struct wrap_of<int16_t> { using type = uint32_t; };
template <> struct wrap_of<uint16_t> { using type = uint32_t; };
template <> struct wrap_of<int32_t> { using type = uint32_t; };
template <> struct wrap_of<int64_t> { using type = uint64_t; };

template <typename T>
__device__ __forceinline__ T sequence_value(size_t i, T start, T step)
{
    if constexpr (std::is_floating_point<T>::value)
    {
        // (T)i * step + start, each rounded (no contraction).
        if constexpr (sizeof(T) == 4)
            return __fadd_rn(__fmul_rn(static_cast<float>(static_cast<uint32_t>(i)), step),
                             start);
        else
            return __dadd_rn(__dmul_rn(static_cast<double>(static_cast<uint32_t>(i)), step),
                             start);
    }
    else
    {
        // (T)i, then the product and sum in the promoted type, converted
        // back: modulo 2^bits of T.
        using W = typename wrap_of<T>::type;
        const W ti = static_cast<W>(static_cast<T>(static_cast<uint32_t>(i)));
        return static_cast<T>(static_cast<W>(static_cast<W>(start)
                                             + ti * static_cast<W>(step)));
    }
}

template <typename T>
__global__ __launch_bounds__(EW_BLOCK) void sequence_kernel(T* __restrict__ out, size_t n,
                                                            T start, T step)
{
    const size_t base = size_t(blockIdx.x) * EW_BLOCK * EW_ITEMS + threadIdx.x;
#pragma unroll
    for (int k = 0; k < EW_ITEMS; ++k)
    {
        const size_t i = base + size_t(k) * EW_BLOCK;
        if (i < n) out[i] = sequence_value<T>(i, start, step);
    }
}

unsigned ew_grid(size_t n)
{
    const size_t per = size_t(EW_BLOCK) * EW_ITEMS;
    return static_cast<unsigned>((n + per - 1) / per);
}

template <typename U> void fill_typed(context& ctx, void* out, size_t n, const void* value)
{
    U v;
    std::memcpy(&v, value, sizeof(U));
    timed_launch tl(ctx, "fill");
    hipLaunchKernelGGL(fill_kernel<U>, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, ctx.stream(),
                       static_cast<U*>(out), n, v);
    CLC_HIP(hipGetLastError());
}

template <typename T>
void sequence_typed(context& ctx, void* out, size_t n, const void* start, const void* step)
{
    T a, b;
    std::memcpy(&a, start, sizeof(T));
    std::memcpy(&b, step, sizeof(T));
    timed_launch tl(ctx, "sequence");
    hipLaunchKernelGGL(sequence_kernel<T>, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, ctx.stream(),
                       static_cast<T*>(out), n, a, b);
    CLC_HIP(hipGetLastError());
}

} // namespace

namespace detail {

void fill(context& ctx, void* out, std::size_t n, int type, const void* value)
{
    const std::size_t es = type_size(type);
    if (value == nullptr) throw std::invalid_argument("clc_fill: null value");
    if (n == 0) return;
    check_size(n);
    if (out == nullptr) throw std::invalid_argument("clc_fill: null buffer");
    ctx.activate();
    switch (es)
    {
        case 1: fill_typed<uint8_t>(ctx, out, n, value); return;
        case 2: fill_typed<uint16_t>(ctx, out, n, value); return;
        case 4: fill_typed<uint32_t>(ctx, out, n, value); return;
        default: fill_typed<uint64_t>(ctx, out, n, value); return;
    }
}

void sequence(context& ctx, void* out, std::size_t n, int type, const void* start,
              const void* step)
{
    type_size(type);
    if (start == nullptr || step == nullptr)
        throw std::invalid_argument("clc_sequence: null start or step");
    if (n == 0) return;
    check_size(n);
    if (out == nullptr) throw std::invalid_argument("clc_sequence: null buffer");
    ctx.activate();
    switch (type)
    {
        case TY_I8: sequence_typed<int8_t>(ctx, out, n, start, step); return;
        case TY_U8: sequence_typed<uint8_t>(ctx, out, n, start, step); return;
        case TY_I16: sequence_typed<int16_t>(ctx, out, n, start, step); return;
        case TY_U16: sequence_typed<uint16_t>(ctx, out, n, start, step); return;
        case TY_I32: sequence_typed<int32_t>(ctx, out, n, start, step); return;
        case TY_U32: sequence_typed<uint32_t>(ctx, out, n, start, step); return;
        case TY_I64: sequence_typed<int64_t>(ctx, out, n, start, step); return;
        case TY_U64: sequence_typed<uint64_t>(ctx, out, n, start, step); return;
        case TY_F32: sequence_typed<float>(ctx, out, n, start, step); return;
        case TY_F64: sequence_typed<double>(ctx, out, n, start, step); return;
    }
}

void transform(context& ctx, const void* in, int in_type, void* out, int out_type,
               std::size_t n, const char* expression)
{
    type_size(in_type);
    type_size(out_type);
    if (expression == nullptr) throw std::invalid_argument("clc_transform: null expression");
    if (n == 0) return;
    check_size(n);
    if (in == nullptr || out == nullptr)
        throw std::invalid_argument("clc_transform: null buffer");
    ctx.activate();
    hipFunction_t f = ctx.programs().function(
        rtc::transform_source(in_type, -1, out_type, expression), rtc::TRANSFORM_KERNEL);
    void* args[] = { &in, &out, &n };
    timed_launch tl(ctx, "rtc_transform");
    CLC_HIP(hipModuleLaunchKernel(f, ew_grid(n), 1, 1, EW_BLOCK, 1, 1, 0, ctx.stream(),
                                  args, nullptr));
}

void transform2(context& ctx, const void* in1, int type1, const void* in2, int type2,
                void* out, int out_type, std::size_t n, const char* expression)
{
    type_size(type1);
    type_size(type2);
    type_size(out_type);
    if (expression == nullptr) throw std::invalid_argument("clc_transform2: null expression");
    if (n == 0) return;
    check_size(n);
    if (in1 == nullptr || in2 == nullptr || out == nullptr)
        throw std::invalid_argument("clc_transform2: null buffer");
    ctx.activate();
    hipFunction_t f = ctx.programs().function(
        rtc::transform_source(type1, type2, out_type, expression), rtc::TRANSFORM2_KERNEL);
    void* args[] = { &in1, &in2, &out, &n };
    timed_launch tl(ctx, "rtc_transform2");
    CLC_HIP(hipModuleLaunchKernel(f, ew_grid(n), 1, 1, EW_BLOCK, 1, 1, 0, ctx.stream(),
                                  args, nullptr));
}

} // namespace detail
} // namespace clc
