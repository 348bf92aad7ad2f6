// Device helpers shared by the gfx950 kernels: wave64 lane arithmetic,
// agent-scope atomics for cross-workgroup status words, and the
// order-preserving key images of the reference sort.
//
// Also compiled at run time by hipRTC (through scan_kernels.hpp), which
// provides the HIP runtime implicitly and has no C library headers.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include <cstddef>
#endif
#include <cstdint>
#include <type_traits>

namespace clc {
namespace dev {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

//! Number of set bits of mask in lanes below the calling lane.
__device__ __forceinline__ uint32_t count_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(
        static_cast<uint32_t>(mask >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Status words are polled by workgroups on other XCDs, whose L2 caches are
// not coherent with each other: agent-scope atomics keep every access at
// the coherence point. Relaxed ordering suffices because each word carries
// its flag and its value together.
template <typename T> __device__ __forceinline__ T load_agent(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename V>
__device__ __forceinline__ void store_agent(T* p, V v)
{
    __hip_atomic_store(p, static_cast<T>(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename V>
__device__ __forceinline__ T fetch_add_agent(T* p, V v)
{
    return __hip_atomic_fetch_add(p, static_cast<T>(v), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

//! Returns *p and stores *p + 1 (wrapping at 2^32), agent scope. Unlike
//! fetch_add_agent from one lane, the compiler keeps it a single atomic
//! whose result is waited for only where it is used.
__device__ __forceinline__ uint32_t inc_agent(uint32_t* p)
{
    return __builtin_amdgcn_atomic_inc32(p, 0xFFFFFFFFu, __ATOMIC_RELAXED, "agent");
}

//! Sum over the wave (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}

//! Inclusive wave64 prefix sum.
__device__ __forceinline__ uint32_t wave_inclusive_add(uint32_t x)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1)
    {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if (lane >= static_cast<uint32_t>(o)) x += y;
    }
    return x;
}

//! Exclusive prefix sum over the first 256 threads of a block of BLOCK >=
//! 256 threads (4 waves); every thread of the block must call it. Threads
//! 256.. pass anything and get garbage. tmp holds 4 words. Ends with a
//! barrier so tmp can be reused.
template <int BLOCK>
__device__ __forceinline__ uint32_t block256_exclusive_add(uint32_t x,
                                                           uint32_t* tmp)
{
    static_assert(BLOCK >= 256, "needs 256 threads");
    const uint32_t tid = threadIdx.x, wave = tid / WAVE;
    const uint32_t inc = wave_inclusive_add(x);
    if (wave < 4 && (tid % WAVE) == WAVE - 1) tmp[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    if (wave >= 1) off += tmp[0];
    if (wave >= 2) off += tmp[1];
    if (wave >= 3) off += tmp[2];
    __syncthreads();
    return off + inc - x;
}

// ---- keys -----------------------------------------------------------------

//! Kinds of key: how the bits map to an order-preserving unsigned image.
enum key_kind : int
{
    KEY_UNSIGNED = 0,
    KEY_SIGNED = 1,
    KEY_FLOAT = 2
};

//! Unsigned image whose ascending order is the sort order: the reference's
//! KEY_TRANSFORM (lib/src/Compute/Algorithm.cpp:459-469) -- uint k; int
//! k ^ sign; float (k & sign) ? ~k : k | sign; descending ~(...) -- and the
//! same formulas at 64 bits.
template <typename K, int KIND, bool DESC>
__device__ __forceinline__ K order_bits(K k)
{
    constexpr int BITS = sizeof(K) * 8;
    constexpr K TOP = K(1) << (BITS - 1);
    K t = k;
    if constexpr (KIND == KEY_SIGNED) t = k ^ TOP;
    if constexpr (KIND == KEY_FLOAT)
    {
        using S = typename std::conditional<sizeof(K) == 4, int32_t,
                                            int64_t>::type;
        const K m = static_cast<K>(static_cast<S>(k) >> (BITS - 1)) | TOP;
        t = k ^ m;
    }
    if constexpr (DESC) t = ~t;
    return t;
}

//! Inverse of order_bits: the key whose image is t.
template <typename K, int KIND, bool DESC>
__device__ __forceinline__ K from_order_bits(K t)
{
    constexpr int BITS = sizeof(K) * 8;
    constexpr K TOP = K(1) << (BITS - 1);
    if constexpr (DESC) t = ~t;
    if constexpr (KIND == KEY_SIGNED) t ^= TOP;
    if constexpr (KIND == KEY_FLOAT) t = (t & TOP) ? (t ^ TOP) : ~t;
    return t;
}

template <typename K, int KIND, bool DESC>
__device__ __forceinline__ uint32_t digit_of(K k, uint32_t shift)
{
    return static_cast<uint32_t>(order_bits<K, KIND, DESC>(k) >> shift) & 0xFFu;
}

} // namespace dev
} // namespace clc
