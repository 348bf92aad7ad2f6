// Prefix scan and reduce for gfx950.
//
// scan: the reference's contract (detail::scan, Algorithm.cpp:444-451;
// kernels :76-127, :280-318): inclusive or exclusive, out may alias in, any
// associative operator including non-commutative ones (operands always
// combine in sequence order), exclusive output seeded with the identity.
// The reference makes three launches per level and recurses over tile
// totals (two full read+write sweeps). Here one launch does it: each tile
// loads 16 bytes per lane per load, reduces its items, publishes its
// aggregate, looks back over earlier tiles' published aggregates/prefixes
// and writes its outputs once (decoupled look-back; 2 x 4 bytes/element of
// HBM traffic for u32).
//
// reduce: the reference's contract (detail::reduce, Algorithm.cpp:400-442):
// op(init, fold) with an order left unspecified (the operator must be
// commutative), n == 0 gives init, and the call blocks until the result is
// on the host. A grid of partial folds, then one workgroup folds the
// partials and applies init.
#include "context.hpp"
#include "device_common.hpp"
#include "ops.hpp"

#include <algorithm>
#include <cstring>

namespace clc {
namespace {

using namespace dev;

// ---- operators ---------------------------------------------------------------

template <typename T> struct wrap_type { using type = T; };
template <> struct wrap_type<int32_t> { using type = uint32_t; };
template <> struct wrap_type<int64_t> { using type = uint64_t; };

//! The canonical operators (reference Algorithm.hpp:53-65) plus the
//! projections. Integer plus/multiply wrap, as in OpenCL C.
template <typename T, int OP>
__device__ __forceinline__ T apply(T a, T b)
{
    using U = typename wrap_type<T>::type;
    if constexpr (OP == OP_PLUS) return static_cast<T>(static_cast<U>(a) + static_cast<U>(b));
    else if constexpr (OP == OP_MULTIPLIES) return static_cast<T>(static_cast<U>(a) * static_cast<U>(b));
    else if constexpr (OP == OP_BIT_AND) return static_cast<T>(a & b);
    else if constexpr (OP == OP_BIT_OR) return static_cast<T>(a | b);
    else if constexpr (OP == OP_BIT_XOR) return static_cast<T>(a ^ b);
    else if constexpr (OP == OP_MINIMUM) return (b < a) ? b : a;
    else if constexpr (OP == OP_MAXIMUM) return (a < b) ? b : a;
    else if constexpr (OP == OP_RIGHT) return b;
    else return a;
}

template <typename T> __device__ __forceinline__ T shfl_up(T x, int o)
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
        uint32_t lo = __shfl_up(static_cast<uint32_t>(u), o, WAVE);
        uint32_t hi = __shfl_up(static_cast<uint32_t>(u >> 32), o, WAVE);
        u = (static_cast<uint64_t>(hi) << 32) | lo;
        T r;
        __builtin_memcpy(&r, &u, 8);
        return r;
    }
}

template <typename T> __device__ __forceinline__ T shfl_down(T x, int o)
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
        uint32_t lo = __shfl_down(static_cast<uint32_t>(u), o, WAVE);
        uint32_t hi = __shfl_down(static_cast<uint32_t>(u >> 32), o, WAVE);
        u = (static_cast<uint64_t>(hi) << 32) | lo;
        T r;
        __builtin_memcpy(&r, &u, 8);
        return r;
    }
}

//! Inclusive wave scan in lane order (lower lanes are left operands).
template <typename T, int OP> __device__ __forceinline__ T wave_scan(T x)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1)
    {
        const T y = shfl_up(x, o);
        if (lane >= static_cast<uint32_t>(o)) x = apply<T, OP>(y, x);
    }
    return x;
}

// ---- scan status words -------------------------------------------------------
// A tile's published value is up to 8 bytes; it is split into two 64-bit
// words, each carrying the flag in its top half, so each word is read and
// written atomically and a reader accepts the pair only when both carry the
// same flag (flags only move 0 -> AGG -> PREFIX).

constexpr uint64_t FLAG_AGG = uint64_t(1) << 32, FLAG_PREFIX = uint64_t(2) << 32;

template <typename T>
__device__ __forceinline__ void publish(uint64_t* st, T v, uint64_t flag)
{
    uint64_t bits = 0;
    __builtin_memcpy(&bits, &v, sizeof(T));
    store_agent(st + 1, flag | (bits >> 32));
    store_agent(st, flag | (bits & uint64_t(0xFFFFFFFFu)));
}

template <typename T, int OP, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void scan_kernel(
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
    for (int j = 1; j < ITEMS; ++j) agg = apply<T, OP>(agg, x[j]);
    const T inc = wave_scan<T, OP>(agg);
    if (lane == WAVE - 1) s_wave[wave] = inc;
    const T excl_lane = shfl_up(inc, 1); // meaningful for lane > 0
    __syncthreads();
    T wave_prefix = s_wave[0]; // meaningful for wave > 0
    for (int w = 1; w < W; ++w)
        if (w < static_cast<int>(wave)) wave_prefix = apply<T, OP>(wave_prefix, s_wave[w]);
    T block_agg = s_wave[0];
    for (int w = 1; w < W; ++w) block_agg = apply<T, OP>(block_agg, s_wave[w]);

    // Look-back by thread 0.
    if (tid == 0)
    {
        uint64_t* mine = status + 2 * tile;
        if (tile == 0)
        {
            publish(mine, block_agg, FLAG_PREFIX);
            s_prefix = identity;
        }
        else
        {
            publish(mine, block_agg, FLAG_AGG);
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
                acc = got ? apply<T, OP>(val, acc) : val;
                got = true;
                if (f == FLAG_PREFIX) break;
                --p;
            }
            publish(mine, apply<T, OP>(acc, block_agg), FLAG_PREFIX);
            s_prefix = acc;
        }
    }
    __syncthreads();

    // pre = tile prefix op wave prefix op lane prefix, each when present.
    T pre = s_prefix;
    bool has = tile > 0;
    if (wave > 0)
    {
        pre = has ? apply<T, OP>(pre, wave_prefix) : wave_prefix;
        has = true;
    }
    if (lane > 0)
    {
        pre = has ? apply<T, OP>(pre, excl_lane) : excl_lane;
        has = true;
    }
    // out = identity op (in[0] op ... op in[i]) (inclusive) or
    //       identity op (in[0] op ... op in[i-1]) (exclusive; identity at 0).
    T y[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
        const T next = has ? apply<T, OP>(pre, x[j]) : x[j];
        y[j] = exclusive ? (has ? apply<T, OP>(identity, pre) : identity)
                         : apply<T, OP>(identity, next);
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

template <typename T, int OP>
__device__ __forceinline__ T block_reduce(T x, T* s_wave)
{
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
    {
        const T y = shfl_down(x, o);
        if (lane + o < WAVE) x = apply<T, OP>(x, y);
    }
    if (lane == 0) s_wave[wave] = x;
    __syncthreads();
    T r = s_wave[0];
    for (uint32_t w = 1; w < blockDim.x / WAVE; ++w) r = apply<T, OP>(r, s_wave[w]);
    return r;
}

template <typename T, int OP, int BLOCK>
__global__ __launch_bounds__(BLOCK) void reduce_kernel(
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
            for (int j = 0; j < VEC; ++j) acc = apply<T, OP>(acc, e[j]);
        }
        start = nvec * VEC;
    }
    for (size_t i = start + size_t(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += stride)
        acc = apply<T, OP>(acc, in[i]);
    const T r = block_reduce<T, OP>(acc, s_wave);
    if (threadIdx.x == 0) out[blockIdx.x] = apply_init ? apply<T, OP>(init, r) : r;
}

// ---- host side -------------------------------------------------------------------

template <typename T, int OP>
void scan_typed(context& ctx, const void* in, void* out, size_t n,
                const void* identity, bool exclusive)
{
    constexpr int BLOCK = 256;
    constexpr int ITEMS = sizeof(T) >= 4 ? 64 / sizeof(T) : 16;
    constexpr size_t TILE = size_t(BLOCK) * ITEMS;
    const size_t tiles = (n + TILE - 1) / TILE;
    T id;
    std::memcpy(&id, identity, sizeof(T));
    const int vec_ok = (reinterpret_cast<uintptr_t>(in) % 16 == 0
                        && reinterpret_cast<uintptr_t>(out) % 16 == 0)
        ? 1 : 0;
    const size_t bytes = 256 + tiles * 16;
    temp_buffer tmp(ctx.resource(), bytes);
    char* t = tmp.as<char>();
    detail::zero_fill(ctx.stream(), t, bytes);
    hipLaunchKernelGGL((scan_kernel<T, OP, BLOCK, ITEMS>), dim3(tiles), dim3(BLOCK),
                       0, ctx.stream(), static_cast<const T*>(in),
                       static_cast<T*>(out), n, id, exclusive ? 1 : 0, vec_ok,
                       reinterpret_cast<uint64_t*>(t + 256),
                       reinterpret_cast<uint32_t*>(t));
    CLC_HIP(hipGetLastError());
}

template <typename T, int OP>
void reduce_typed(context& ctx, const void* in, size_t n, const void* identity,
                  const void* init, void* result)
{
    constexpr int BLOCK = 256;
    T id, ini;
    std::memcpy(&id, identity, sizeof(T));
    std::memcpy(&ini, init, sizeof(T));
    const size_t per_block = size_t(BLOCK) * 16;
    const uint32_t groups = static_cast<uint32_t>(std::max<size_t>(
        1, std::min<size_t>((n + per_block - 1) / per_block,
                            size_t(ctx.facts().cu_count) * 4)));
    temp_buffer partials(ctx.resource(), (groups + 1) * sizeof(T));
    T* part = partials.as<T>();
    const int vec_ok = reinterpret_cast<uintptr_t>(in) % 16 == 0 ? 1 : 0;
    hipLaunchKernelGGL((reduce_kernel<T, OP, BLOCK>), dim3(groups), dim3(BLOCK), 0,
                       ctx.stream(), static_cast<const T*>(in), n, id, vec_ok,
                       part, 0, ini);
    hipLaunchKernelGGL((reduce_kernel<T, OP, BLOCK>), dim3(1), dim3(BLOCK), 0,
                       ctx.stream(), part, size_t(groups), id, 0, part + groups,
                       1, ini);
    CLC_HIP(hipGetLastError());
    CLC_HIP(hipMemcpyAsync(ctx.pinned(), part + groups, sizeof(T),
                           hipMemcpyDeviceToHost, ctx.stream()));
    CLC_HIP(hipStreamSynchronize(ctx.stream()));
    std::memcpy(result, ctx.pinned(), sizeof(T));
}

template <template <typename, int> class F, typename T, typename... A>
void by_op_int(int op, A&&... a)
{
    switch (op)
    {
        case OP_PLUS: F<T, OP_PLUS>::run(a...); return;
        case OP_MULTIPLIES: F<T, OP_MULTIPLIES>::run(a...); return;
        case OP_BIT_AND: F<T, OP_BIT_AND>::run(a...); return;
        case OP_BIT_OR: F<T, OP_BIT_OR>::run(a...); return;
        case OP_BIT_XOR: F<T, OP_BIT_XOR>::run(a...); return;
        case OP_MINIMUM: F<T, OP_MINIMUM>::run(a...); return;
        case OP_MAXIMUM: F<T, OP_MAXIMUM>::run(a...); return;
        case OP_RIGHT: F<T, OP_RIGHT>::run(a...); return;
        case OP_LEFT: F<T, OP_LEFT>::run(a...); return;
    }
    throw std::invalid_argument("clc: unknown operator");
}

template <template <typename, int> class F, typename T, typename... A>
void by_op_float(int op, A&&... a)
{
    switch (op)
    {
        case OP_PLUS: F<T, OP_PLUS>::run(a...); return;
        case OP_MULTIPLIES: F<T, OP_MULTIPLIES>::run(a...); return;
        case OP_MINIMUM: F<T, OP_MINIMUM>::run(a...); return;
        case OP_MAXIMUM: F<T, OP_MAXIMUM>::run(a...); return;
        case OP_RIGHT: F<T, OP_RIGHT>::run(a...); return;
        case OP_LEFT: F<T, OP_LEFT>::run(a...); return;
    }
    throw std::invalid_argument("clc: operator undefined for floating-point elements");
}

template <template <typename, int> class F, typename... A>
void by_type(int type, int op, A&&... a)
{
    switch (type)
    {
        case TY_I8: by_op_int<F, int8_t>(op, a...); return;
        case TY_U8: by_op_int<F, uint8_t>(op, a...); return;
        case TY_I16: by_op_int<F, int16_t>(op, a...); return;
        case TY_U16: by_op_int<F, uint16_t>(op, a...); return;
        case TY_I32: by_op_int<F, int32_t>(op, a...); return;
        case TY_U32: by_op_int<F, uint32_t>(op, a...); return;
        case TY_I64: by_op_int<F, int64_t>(op, a...); return;
        case TY_U64: by_op_int<F, uint64_t>(op, a...); return;
        case TY_F32: by_op_float<F, float>(op, a...); return;
        case TY_F64: by_op_float<F, double>(op, a...); return;
    }
    throw std::invalid_argument("clc: unknown element type");
}

template <typename T, int OP> struct scan_fn
{
    static void run(context& ctx, const void* in, void* out, size_t n,
                    const void* id, bool ex)
    {
        scan_typed<T, OP>(ctx, in, out, n, id, ex);
    }
};

template <typename T, int OP> struct reduce_fn
{
    static void run(context& ctx, const void* in, size_t n, const void* id,
                    const void* init, void* result)
    {
        reduce_typed<T, OP>(ctx, in, n, id, init, result);
    }
};

} // namespace

namespace detail {

void scan(context& ctx, const void* in, void* out, std::size_t n, int type,
          int op, const void* identity, bool exclusive)
{
    if (n == 0) return;
    check_size(n);
    if (in == nullptr || out == nullptr)
        throw std::invalid_argument("clc_scan: null buffer");
    ctx.activate();
    by_type<scan_fn>(type, op, ctx, in, out, n, identity, exclusive);
}

void reduce(context& ctx, const void* in, std::size_t n, int type, int op,
            const void* identity, const void* init, void* result)
{
    if (n == 0)
    {
        std::memcpy(result, init, type_size(type));
        return;
    }
    check_size(n);
    if (in == nullptr) throw std::invalid_argument("clc_reduce: null buffer");
    ctx.activate();
    by_type<reduce_fn>(type, op, ctx, in, n, identity, init, result);
}

} // namespace detail
} // namespace clc
