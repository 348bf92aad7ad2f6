// Scan, reduce and copy_if kernel bodies for gfx950, templated on the
// element type T and an operator functor Op (static __device__ T apply(T a,
// T b)) or a predicate functor Pred (static __device__ bool test(T x)).
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

//! Threads per scan workgroup and per reduce workgroup.
constexpr int SCAN_BLOCK = 512;
constexpr int REDUCE_BLOCK = 256;

//! Elements per thread of a scan tile for elements of `bytes` bytes: 128
//! bytes a thread (eight 16-byte loads) for 4- and 8-byte elements, 32
//! elements for narrower ones. Tiles of 16K+ elements (64 KiB of u32) keep
//! the look-back's latency per tile small against the tile's traffic: on
//! 2^28 u32, 4K-element tiles ran at half the speed of 16K-element ones.
constexpr int scan_items_of(unsigned bytes) { return bytes >= 4 ? int(128 / bytes) : 32; }
template <typename T> constexpr int scan_items() { return scan_items_of(sizeof(T)); }

// Streaming 16-byte accesses: each element is read once and written once,
// so loads and stores carry the non-temporal hint (a copy of 2^28 u32 with
// the scan's access pattern: 0.39 ms plain, 0.36 ms non-temporal).
typedef unsigned int stream_v4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 load_stream(const uint4* p)
{
    const stream_v4 v = __builtin_nontemporal_load(reinterpret_cast<const stream_v4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void store_stream(uint4* p, uint4 q)
{
    stream_v4 v;
    v.x = q.x;
    v.y = q.y;
    v.z = q.z;
    v.w = q.w;
    __builtin_nontemporal_store(v, reinterpret_cast<stream_v4*>(p));
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
// Each tile publishes its aggregate, then its inclusive prefix, with a flag
// in the top half of 64-bit words read and written atomically. Elements of
// up to 4 bytes share one word with the flag; 8-byte elements are split
// over two words, each carrying the flag, and a reader accepts the pair only
// when both carry the same flag (flags only move 0 -> AGG -> PREFIX).

constexpr uint64_t SCAN_FLAG_AGG = uint64_t(1) << 32;
constexpr uint64_t SCAN_FLAG_PREFIX = uint64_t(2) << 32;
constexpr uint64_t SCAN_LOW = 0xFFFFFFFFu;

template <typename T>
__device__ __forceinline__ void publish_status(uint64_t* st, T v, uint64_t flag)
{
    uint64_t bits = 0;
    __builtin_memcpy(&bits, &v, sizeof(T));
    if constexpr (sizeof(T) > 4) store_agent(st + 1, flag | (bits >> 32));
    store_agent(st, flag | (bits & SCAN_LOW));
}

//! Reads tile p's status: its flag (0 while not ready) and value bits.
template <typename T>
__device__ __forceinline__ uint64_t read_status(const uint64_t* status, long long p,
                                                uint64_t& bits)
{
    const uint64_t lo = load_agent(status + 2 * p);
    uint64_t f = lo & ~SCAN_LOW;
    bits = lo & SCAN_LOW;
    if constexpr (sizeof(T) > 4)
    {
        const uint64_t hi = load_agent(status + 2 * p + 1);
        if ((hi & ~SCAN_LOW) != f) f = 0; // torn pair: not ready yet
        bits |= hi << 32;
    }
    return f;
}

//! Combination, earliest first, of the values published by the tiles
//! before tile (tile > 0), by one wave: each round lanes 0..WIN-1 read the
//! status words of WIN consecutive predecessors at once (lane l: tile - 1 -
//! l - WIN*r). A round ends the walk when some lane holds a prefix and every
//! lane before it an aggregate; when all WIN hold aggregates they are folded
//! in and the walk goes on; otherwise the round reads the same predecessors
//! again. Lanes are combined by an ordered tree (a higher lane is an earlier
//! tile, so it is the left operand). The result is valid in lane 0.
template <typename T, typename Op, int WIN>
__device__ __forceinline__ T window_lookback(const uint64_t* status, size_t tile)
{
    static_assert(WIN >= 1 && WIN <= WAVE && (WIN & (WIN - 1)) == 0, "window");
    const uint32_t lane = lane_id();
    const bool reader = lane < static_cast<uint32_t>(WIN);
    T acc{};
    bool got = false;
    long long base = static_cast<long long>(tile) - 1;
    for (;;)
    {
        const long long p = base - static_cast<long long>(lane);
        uint64_t bits = 0, f = SCAN_FLAG_PREFIX; // before tile 0: ends the walk
        if (reader && p >= 0) f = read_status<T>(status, p, bits);
        const uint64_t prefix = __ballot(reader && f == SCAN_FLAG_PREFIX);
        const uint64_t waiting = __ballot(reader && f == 0);
        // Lanes 0..last are needed: up to the first prefix, or all WIN.
        const int last = prefix ? __builtin_ctzll(prefix) : WIN - 1;
        const uint64_t needed = (uint64_t(2) << last) - 1; // last < 64
        if (waiting & needed) continue;
        T x;
        __builtin_memcpy(&x, &bits, sizeof(T));
#pragma unroll
        for (int o = 1; o < WIN; o <<= 1)
        {
            const T y = shfl_down_elem(x, o);
            if ((lane & (2 * o - 1)) == 0 && static_cast<int>(lane) + o <= last)
                x = Op::apply(y, x);
        }
        acc = got ? Op::apply(x, acc) : x;
        got = true;
        if (prefix) return acc;
        base -= WIN;
    }
}

//! Value of lane src of the wave, in every lane.
template <typename T> __device__ __forceinline__ T lane_value(T x, int src)
{
    if constexpr (sizeof(T) <= 4)
    {
        uint32_t u = 0;
        __builtin_memcpy(&u, &x, sizeof(T));
        u = __builtin_amdgcn_readlane(u, src);
        T r;
        __builtin_memcpy(&r, &u, sizeof(T));
        return r;
    }
    else
    {
        uint64_t u;
        __builtin_memcpy(&u, &x, 8);
        const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), src);
        const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), src);
        u = (static_cast<uint64_t>(hi) << 32) | lo;
        T r;
        __builtin_memcpy(&r, &u, 8);
        return r;
    }
}

//! One tile of a single-pass scan (decoupled look-back): the body of a
//! kernel launched with one BLOCK-thread workgroup per tile of BLOCK*ITEMS
//! elements. status holds 2 zeroed words per tile, tile_counter one zeroed
//! word (tiles are numbered in the order workgroups start, so every tile a
//! workgroup waits on belongs to a workgroup already running).
//!
//! Layout: wave w owns the w-th span of ITEMS*64 consecutive elements and
//! reads it with J = ITEMS/VEC fully coalesced 16-byte loads (load j, lane l:
//! elements j*64*VEC + l*VEC .. +VEC-1 of the span). Each load's VEC
//! elements are folded in the thread, the folds scanned across the wave,
//! the J span pieces chained in order, then the W spans in order, then the
//! tile prefix from the look-back is put in front.
template <typename T, typename Op, int BLOCK, int ITEMS, int WIN = 16>
__device__ __forceinline__ void scan_tile(
    const T* in, T* out, size_t n, T identity, int exclusive, int vec_ok,
    uint64_t* status, uint32_t* tile_counter)
{
    constexpr int W = BLOCK / WAVE;
    constexpr int VEC = 16 / sizeof(T); // elements per 16-byte load
    static_assert(ITEMS % VEC == 0, "items per thread");
    constexpr int J = ITEMS / VEC;
    constexpr size_t SPAN = size_t(WAVE) * ITEMS;
    constexpr size_t TILE = size_t(BLOCK) * ITEMS;
    __shared__ T s_wave[W];
    __shared__ T s_prefix;
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    __syncthreads();
    const size_t tile = s_tile;
    const size_t first = tile * TILE + wave * SPAN + size_t(lane) * VEC;
    const bool full = (tile + 1) * TILE <= n && vec_ok;

    T x[ITEMS];
    if (full)
    {
        const uint4* src = reinterpret_cast<const uint4*>(in + first);
#pragma unroll
        for (int j = 0; j < J; ++j)
        {
            const uint4 q = load_stream(src + j * WAVE);
            __builtin_memcpy(&x[j * VEC], &q, 16);
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int e = 0; e < VEC; ++e)
            {
                const size_t i = first + size_t(j) * WAVE * VEC + e;
                x[j * VEC + e] = i < n ? in[i] : identity;
            }
    }

    // Prefixes are computed unseeded (no identity inside), so operators
    // without a two-sided identity (the projections) still give the
    // sequential result; the identity is applied once, on output.
    T inc[J];  // inclusive scan of load j's folds across the wave
    T span{};  // the wave's span aggregate (wave-uniform)
#pragma unroll
    for (int j = 0; j < J; ++j)
    {
        T f = x[j * VEC];
#pragma unroll
        for (int e = 1; e < VEC; ++e) f = Op::apply(f, x[j * VEC + e]);
        inc[j] = wave_scan_op<T, Op>(f);
        const T total = lane_value(inc[j], WAVE - 1);
        span = j == 0 ? total : Op::apply(span, total);
    }
    if (lane == 0) s_wave[wave] = span;
    __syncthreads();
    T wave_prefix = s_wave[0]; // meaningful for wave > 0
    for (int w = 1; w < W; ++w)
        if (w < static_cast<int>(wave)) wave_prefix = Op::apply(wave_prefix, s_wave[w]);
    T block_agg = s_wave[0];
    for (int w = 1; w < W; ++w) block_agg = Op::apply(block_agg, s_wave[w]);

    if (wave == 0)
    {
        uint64_t* mine = status + 2 * tile;
        if (tile == 0)
        {
            if (lane == 0)
            {
                publish_status(mine, block_agg, SCAN_FLAG_PREFIX);
                s_prefix = identity;
            }
        }
        else
        {
            if (lane == 0) publish_status(mine, block_agg, SCAN_FLAG_AGG);
            const T acc = window_lookback<T, Op, WIN>(status, tile);
            if (lane == 0)
            {
                publish_status(mine, Op::apply(acc, block_agg), SCAN_FLAG_PREFIX);
                s_prefix = acc;
            }
        }
    }
    __syncthreads();

    // run = tile prefix op earlier spans op earlier pieces of this span,
    // each when present (wave-uniform).
    T run = s_prefix;
    bool has = tile > 0;
    if (wave > 0)
    {
        run = has ? Op::apply(run, wave_prefix) : wave_prefix;
        has = true;
    }
    // out = identity op (in[0] op ... op in[i]) (inclusive) or
    //       identity op (in[0] op ... op in[i-1]) (exclusive; identity at 0).
    T y[ITEMS];
#pragma unroll
    for (int j = 0; j < J; ++j)
    {
        const T excl = shfl_up_elem(inc[j], 1); // lanes below, lane > 0
        T pre = run;
        bool hp = has;
        if (lane > 0)
        {
            pre = hp ? Op::apply(pre, excl) : excl;
            hp = true;
        }
#pragma unroll
        for (int e = 0; e < VEC; ++e)
        {
            const T v = x[j * VEC + e];
            const T next = hp ? Op::apply(pre, v) : v;
            y[j * VEC + e] = exclusive ? (hp ? Op::apply(identity, pre) : identity)
                                       : Op::apply(identity, next);
            pre = next;
            hp = true;
        }
        const T total = lane_value(inc[j], WAVE - 1);
        run = has ? Op::apply(run, total) : total;
        has = true;
    }
    if (full)
    {
        uint4* dst = reinterpret_cast<uint4*>(out + first);
#pragma unroll
        for (int j = 0; j < J; ++j)
        {
            uint4 q;
            __builtin_memcpy(&q, &y[j * VEC], 16);
            store_stream(dst + j * WAVE, q);
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int e = 0; e < VEC; ++e)
            {
                const size_t i = first + size_t(j) * WAVE * VEC + e;
                if (i < n) out[i] = y[j * VEC + e];
            }
    }
}

// ---- copy_if (stream compaction) -----------------------------------------------

struct count_plus
{
    static __device__ __forceinline__ uint32_t apply(uint32_t a, uint32_t b) { return a + b; }
};

//! One tile of a single-pass, order-preserving copy_if: the elements x of
//! the tile with Pred::test(x) are written to out at their rank among all
//! such elements of in[0, n). Same tiles, layout, status words and
//! look-back as scan_tile, over kept-element counts. Each span piece ranks
//! its kept elements with one ballot per element position and the wave
//! packs them into its LDS staging row before the look-back; after it, the
//! wave writes the row to out as one contiguous run (on 2^28 u32, half
//! kept: 0.37 ms staged against 0.45 ms storing each element at its rank
//! from registers, which also held 107 VGPRs). The last tile writes the
//! total count to *count. out may alias in: an element moves only to a
//! lower or equal index, after every earlier tile has read its input (a
//! tile's aggregate is published after its loads).
template <typename T, typename Pred, int BLOCK, int ITEMS, int WIN = 16>
__device__ __forceinline__ void select_tile(
    const T* in, T* out, size_t n, int vec_ok, uint64_t* status,
    uint32_t* tile_counter, uint32_t* count)
{
    constexpr int W = BLOCK / WAVE;
    constexpr int VEC = 16 / sizeof(T);
    static_assert(ITEMS % VEC == 0, "items per thread");
    constexpr int J = ITEMS / VEC;
    constexpr size_t SPAN = size_t(WAVE) * ITEMS;
    constexpr size_t TILE = size_t(BLOCK) * ITEMS;
    __shared__ T s_stage[W][SPAN];
    __shared__ uint32_t s_wave[W];
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    __syncthreads();
    const size_t tile = s_tile;
    const size_t first = tile * TILE + wave * SPAN + size_t(lane) * VEC;
    const bool full = (tile + 1) * TILE <= n;

    T x[ITEMS];
    if (full && vec_ok)
    {
        const uint4* src = reinterpret_cast<const uint4*>(in + first);
#pragma unroll
        for (int j = 0; j < J; ++j)
        {
            const uint4 q = load_stream(src + j * WAVE);
            __builtin_memcpy(&x[j * VEC], &q, 16);
        }
    }
    else
    {
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int e = 0; e < VEC; ++e)
            {
                const size_t i = first + size_t(j) * WAVE * VEC + e;
                x[j * VEC + e] = (full || i < n) ? in[i] : T{};
            }
    }

    // Pack the span's kept elements, in order, into the wave's LDS row.
    uint32_t kept = 0; // wave-uniform
    const uint64_t below = (uint64_t(1) << lane) - 1;
#pragma unroll
    for (int j = 0; j < J; ++j)
    {
        bool f[VEC];
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int e = 0; e < VEC; ++e)
        {
            const size_t i = first + size_t(j) * WAVE * VEC + e;
            f[e] = (full || i < n) && Pred::test(x[j * VEC + e]);
            const uint64_t m = __ballot(f[e]);
            before += __builtin_popcountll(m & below);
            total += __builtin_popcountll(m);
        }
        uint32_t own = kept + before;
#pragma unroll
        for (int e = 0; e < VEC; ++e)
            if (f[e]) s_stage[wave][own++] = x[j * VEC + e];
        kept += total;
    }
    if (lane == 0) s_wave[wave] = kept;
    __syncthreads();
    uint32_t wave_prefix = 0, block_total = 0;
    for (int w = 0; w < W; ++w)
    {
        if (w < static_cast<int>(wave)) wave_prefix += s_wave[w];
        block_total += s_wave[w];
    }

    if (wave == 0)
    {
        uint64_t* mine = status + 2 * tile;
        uint32_t prefix = 0;
        if (tile == 0)
        {
            if (lane == 0) publish_status(mine, block_total, SCAN_FLAG_PREFIX);
        }
        else
        {
            if (lane == 0) publish_status(mine, block_total, SCAN_FLAG_AGG);
            prefix = window_lookback<uint32_t, count_plus, WIN>(status, tile);
            if (lane == 0) publish_status(mine, prefix + block_total, SCAN_FLAG_PREFIX);
        }
        if (lane == 0)
        {
            s_prefix = prefix;
            if ((tile + 1) * TILE >= n) *count = prefix + block_total;
        }
    }
    __syncthreads();

    T* dst = out + size_t(s_prefix) + wave_prefix;
    for (uint32_t k = lane; k < kept; k += WAVE)
        __builtin_nontemporal_store(s_stage[wave][k], dst + k);
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
        // One 16-byte load a thread per grid stride: on 2^28 u32 this ran at
        // 6.4 TB/s with 8 workgroups per CU; four loads in flight a thread
        // (grid-strided or block-contiguous) ran at 5.2-6.2 TB/s.
        const size_t nvec = n / VEC;
        const uint4* v = reinterpret_cast<const uint4*>(in);
        size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x;
        for (; i < nvec; i += stride)
        {
            const uint4 q = load_stream(v + i);
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
