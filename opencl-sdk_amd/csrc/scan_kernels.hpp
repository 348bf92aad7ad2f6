// Scan and reduce kernel bodies for gfx950, templated on the element type T
// and an operator functor Op (static __device__ T apply(T a, T b)).
//
// The same bodies serve the canonical operators, compiled ahead of time
// (scan.hip), and any other operator expression, compiled at run time by
// hipRTC with the expression pasted into a functor (rtc.cpp), as the
// reference pastes OP and IDENTITY into its OpenCL C scan and reduce sources
// (lib/src/Compute/Algorithm.cpp:40-127, :280-318, :400-442). This header
// and device_common.hpp are therefore self-contained: hipRTC provides the
// HIP runtime implicitly and has no C library headers.
#pragma once

#include "device_common.hpp"

namespace clc {
namespace dev {

//! Threads per scan / reduce workgroup.
constexpr int SCAN_BLOCK = 256;

//! Elements per thread of a scan tile: 64 bytes per thread for 4- and
//! 8-byte elements (four 16-byte loads), 16 elements for narrower ones.
template <typename T> constexpr int scan_items()
{
    return sizeof(T) >= 4 ? int(64 / sizeof(T)) : 16;
}

template <typename T> __device__ __forceinline__ T shfl_up_elem(T x, int o)
{
    if constexpr (sizeof(T) <= 4)
    {
        uint32_t u = 0;
        __builtin_memcpy(&u, &x, sizeof(T));
        u = __shfl_up(u, o, WAVE);
        T r;
        __builtin_memcpy(&r, &u, sizeof(T));
        return r;
    }
    else
    {
        uint64_t u;
        __builtin_memcpy(&u, &x, 8);
        const uint32_t lo = __shfl_up(static_cast<uint32_t>(u), o, WAVE);
        const uint32_t hi = __shfl_up(static_cast<uint32_t>(u >> 32), o, WAVE);
        u = (static_cast<uint64_t>(hi) << 32) | lo;
        T r;
        __builtin_memcpy(&r, &u, 8);
        return r;
    }
}

template <typename T> __device__ __forceinline__ T shfl_down_elem(T x, int o)
{
    if constexpr (sizeof(T) <= 4)
    {
        uint32_t u = 0;
        __builtin_memcpy(&u, &x, sizeof(T));
        u = __shfl_down(u, o, WAVE);
        T r;
        __builtin_memcpy(&r, &u, sizeof(T));
        return r;
    }
    else
    {
        uint64_t u;
        __builtin_memcpy(&u, &x, 8);
        const uint32_t lo = __shfl_down(static_cast<uint32_t>(u), o, WAVE);
        const uint32_t hi = __shfl_down(static_cast<uint32_t>(u >> 32), o, WAVE);
        u = (static_cast<uint64_t>(hi) << 32) | lo;
        T r;
        __builtin_memcpy(&r, &u, 8);
        return r;
    }
}

//! Inclusive wave scan in lane order (lower lanes are left operands).
template <typename T, typename Op> __device__ __forceinline__ T wave_scan_op(T x)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1)
    {
        const T y = shfl_up_elem(x, o);
        if (lane >= static_cast<uint32_t>(o)) x = Op::apply(y, x);
    }
    return x;
}

// ---- scan status words -------------------------------------------------------
// A tile's published value is up to 8 bytes; it is split into two 64-bit
// words, each carrying the flag in its top half, so each word is read and
// written atomically and a reader accepts the pair only when both carry the
// same flag (flags only move 0 -> AGG -> PREFIX).

constexpr uint64_t SCAN_FLAG_AGG = uint64_t(1) << 32;
constexpr uint64_t SCAN_FLAG_PREFIX = uint64_t(2) << 32;

template <typename T>
__device__ __forceinline__ void publish_status(uint64_t* st, T v, uint64_t flag)
{
    uint64_t bits = 0;
    __builtin_memcpy(&bits, &v, sizeof(T));
    store_agent(st + 1, flag | (bits >> 32));
    store_agent(st, flag | (bits & uint64_t(0xFFFFFFFFu)));
}

//! One tile of a single-pass scan (decoupled look-back): the body of a
//! kernel launched with one BLOCK-thread workgroup per tile of BLOCK*ITEMS
//! elements. status holds 2 zeroed words per tile, tile_counter one zeroed
//! word (tiles are numbered in the order workgroups start, so every tile a
//! workgroup waits on belongs to a workgroup already running).
template <typename T, typename Op, int BLOCK, int ITEMS>
__device__ __forceinline__ void scan_tile(
    const T* in, T* out, size_t n, T identity, int exclusive, int vec_ok,
    uint64_t* status, uint32_t* tile_counter)
{
    constexpr int W = BLOCK / WAVE;
    constexpr size_t TILE = size_t(BLOCK) * ITEMS;
    constexpr int VEC = 16 / sizeof(T); // elements per 16-byte load
    static_assert(ITEMS % VEC == 0, "items per thread");
    __shared__ T s_wave[W];
    __shared__ T s_prefix;
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    __syncthreads();
    const size_t tile = s_tile;
    const size_t first = tile * TILE + size_t(tid) * ITEMS;
    const bool full = (tile + 1) * TILE <= n && vec_ok;

    T x[ITEMS];
    if (full)
    {
        const uint4* src = reinterpret_cast<const uint4*>(in + first);
#pragma unroll
        for (int j = 0; j < ITEMS / VEC; ++j)
        {
            const uint4 q = src[j];
            __builtin_memcpy(&x[j * VEC], &q, 16);
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            x[j] = first + j < n ? in[first + j] : identity;
    }

    // Prefixes are computed unseeded (no identity inside), so operators
    // without a two-sided identity (the projections) still give the
    // sequential result; the identity is applied once, on output.
    // Thread fold, then wave and block exclusive scans of the folds.
    T agg = x[0];
#pragma unroll
    for (int j = 1; j < ITEMS; ++j) agg = Op::apply(agg, x[j]);
    const T inc = wave_scan_op<T, Op>(agg);
    if (lane == WAVE - 1) s_wave[wave] = inc;
    const T excl_lane = shfl_up_elem(inc, 1); // meaningful for lane > 0
    __syncthreads();
    T wave_prefix = s_wave[0]; // meaningful for wave > 0
    for (int w = 1; w < W; ++w)
        if (w < static_cast<int>(wave)) wave_prefix = Op::apply(wave_prefix, s_wave[w]);
    T block_agg = s_wave[0];
    for (int w = 1; w < W; ++w) block_agg = Op::apply(block_agg, s_wave[w]);

    // Look-back by thread 0.
    if (tid == 0)
    {
        uint64_t* mine = status + 2 * tile;
        if (tile == 0)
        {
            publish_status(mine, block_agg, SCAN_FLAG_PREFIX);
            s_prefix = identity;
        }
        else
        {
            publish_status(mine, block_agg, SCAN_FLAG_AGG);
            T acc = identity;
            bool got = false;
            size_t p = tile - 1;
            for (;;)
            {
                const uint64_t lo = load_agent(status + 2 * p);
                const uint64_t hi = load_agent(status + 2 * p + 1);
                const uint64_t f = lo & ~uint64_t(0xFFFFFFFFu);
                if (f == 0 || f != (hi & ~uint64_t(0xFFFFFFFFu))) continue;
                const uint64_t bits = (lo & uint64_t(0xFFFFFFFFu)) | (hi << 32);
                T val;
                __builtin_memcpy(&val, &bits, sizeof(T));
                acc = got ? Op::apply(val, acc) : val;
                got = true;
                if (f == SCAN_FLAG_PREFIX) break;
                --p;
            }
            publish_status(mine, Op::apply(acc, block_agg), SCAN_FLAG_PREFIX);
            s_prefix = acc;
        }
    }
    __syncthreads();

    // pre = tile prefix op wave prefix op lane prefix, each when present.
    T pre = s_prefix;
    bool has = tile > 0;
    if (wave > 0)
    {
        pre = has ? Op::apply(pre, wave_prefix) : wave_prefix;
        has = true;
    }
    if (lane > 0)
    {
        pre = has ? Op::apply(pre, excl_lane) : excl_lane;
        has = true;
    }
    // out = identity op (in[0] op ... op in[i]) (inclusive) or
    //       identity op (in[0] op ... op in[i-1]) (exclusive; identity at 0).
    T y[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
        const T next = has ? Op::apply(pre, x[j]) : x[j];
        y[j] = exclusive ? (has ? Op::apply(identity, pre) : identity)
                         : Op::apply(identity, next);
        pre = next;
        has = true;
    }
    if (full)
    {
        uint4* dst = reinterpret_cast<uint4*>(out + first);
#pragma unroll
        for (int j = 0; j < ITEMS / VEC; ++j)
        {
            uint4 q;
            __builtin_memcpy(&q, &y[j * VEC], 16);
            dst[j] = q;
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            if (first + j < n) out[first + j] = y[j];
    }
}

// ---- reduce --------------------------------------------------------------------

template <typename T, typename Op>
__device__ __forceinline__ T block_reduce_op(T x, T* s_wave)
{
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
    {
        const T y = shfl_down_elem(x, o);
        if (lane + o < WAVE) x = Op::apply(x, y);
    }
    if (lane == 0) s_wave[wave] = x;
    __syncthreads();
    T r = s_wave[0];
    for (uint32_t w = 1; w < blockDim.x / WAVE; ++w) r = Op::apply(r, s_wave[w]);
    return r;
}

//! One workgroup's fold of a grid-strided share of in[0, n): the body of
//! the reduce kernel. out[blockIdx.x] = fold, or op(init, fold) when
//! apply_init (the second launch, one workgroup over the partials).
template <typename T, typename Op, int BLOCK>
__device__ __forceinline__ void reduce_partial(
    const T* __restrict__ in, size_t n, T identity, int vec_ok,
    T* __restrict__ out, int apply_init, T init)
{
    __shared__ T s_wave[BLOCK / WAVE];
    constexpr int VEC = 16 / sizeof(T);
    T acc = identity;
    const size_t stride = size_t(gridDim.x) * BLOCK;
    size_t start = 0;
    if (vec_ok)
    {
        const size_t nvec = n / VEC;
        const uint4* v = reinterpret_cast<const uint4*>(in);
        for (size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nvec; i += stride)
        {
            const uint4 q = v[i];
            T e[VEC];
            __builtin_memcpy(e, &q, 16);
#pragma unroll
            for (int j = 0; j < VEC; ++j) acc = Op::apply(acc, e[j]);
        }
        start = nvec * VEC;
    }
    for (size_t i = start + size_t(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += stride)
        acc = Op::apply(acc, in[i]);
    const T r = block_reduce_op<T, Op>(acc, s_wave);
    if (threadIdx.x == 0) out[blockIdx.x] = apply_init ? Op::apply(init, r) : r;
}

} // namespace dev
} // namespace clc
