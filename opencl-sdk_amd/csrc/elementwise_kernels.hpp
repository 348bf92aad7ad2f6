// Elementwise kernel bodies for gfx950: transform (out[i] = F(x = in[i]))
// and transform2 (out[i] = F(x = in1[i], y = in2[i])), templated on the
// element types and an expression functor. Compiled at run time by hipRTC
// with the caller's expression (rtc.cpp), as the reference pastes EXPR into
// its OpenCL C transform sources (lib/src/Compute/Algorithm.cpp:30-56,
// :361-398); self-contained for that reason, like scan_kernels.hpp.
//
// Layout: a workgroup of BLOCK threads handles BLOCK * ITEMS consecutive
// elements, item k of thread t at k * BLOCK + t, so every load and store
// instruction of a wave covers 64 consecutive elements. All loads of a
// thread are issued before its stores; full tiles stream (non-temporal
// loads and stores: 2^28 u32 "x * 3u + 1u" 0.375 -> 0.364 ms).
#pragma once

#include "device_common.hpp"

namespace clc {
namespace dev {

constexpr int EW_BLOCK = 256;
constexpr int EW_ITEMS = 8;

template <typename TI, typename TO, typename F, int BLOCK, int ITEMS>
__device__ __forceinline__ void transform_tile(const TI* __restrict__ in,
                                               TO* __restrict__ out, size_t n)
{
    const size_t base = size_t(blockIdx.x) * BLOCK * ITEMS + threadIdx.x;
    TI x[ITEMS];
    if (base + size_t(ITEMS - 1) * BLOCK < n)
    {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            x[k] = __builtin_nontemporal_load(in + base + size_t(k) * BLOCK);
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            __builtin_nontemporal_store(F::apply(x[k]), out + base + size_t(k) * BLOCK);
    }
    else
    {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
        {
            const size_t i = base + size_t(k) * BLOCK;
            if (i < n) out[i] = F::apply(in[i]);
        }
    }
}

template <typename T1, typename T2, typename TO, typename F, int BLOCK, int ITEMS>
__device__ __forceinline__ void transform2_tile(const T1* __restrict__ in1,
                                                const T2* __restrict__ in2,
                                                TO* __restrict__ out, size_t n)
{
    const size_t base = size_t(blockIdx.x) * BLOCK * ITEMS + threadIdx.x;
    T1 x[ITEMS];
    T2 y[ITEMS];
    if (base + size_t(ITEMS - 1) * BLOCK < n)
    {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
        {
            x[k] = __builtin_nontemporal_load(in1 + base + size_t(k) * BLOCK);
            y[k] = __builtin_nontemporal_load(in2 + base + size_t(k) * BLOCK);
        }
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            __builtin_nontemporal_store(F::apply(x[k], y[k]), out + base + size_t(k) * BLOCK);
    }
    else
    {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
        {
            const size_t i = base + size_t(k) * BLOCK;
            if (i < n) out[i] = F::apply(in1[i], in2[i]);
        }
    }
}

} // namespace dev
} // namespace clc
