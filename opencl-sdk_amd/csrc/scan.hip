// Prefix scan and reduce for gfx950.
//
// scan: the reference's contract (detail::scan, Algorithm.cpp:444-451;
// kernels :76-127, :280-318): inclusive or exclusive, out may alias in, any
// associative operator including non-commutative ones (operands always
// combine in sequence order), exclusive output seeded with the identity.
// The reference makes three launches per level and recurses over tile
// totals (two full read+write sweeps). Here one launch does it: each
// 512-thread tile reads its 64 KiB (u32) with coalesced 16-byte loads,
// reduces them, publishes its aggregate, looks back over 16 earlier tiles'
// published aggregates/prefixes at a time and writes its outputs once
// (decoupled look-back; 2 x 4 bytes/element of HBM traffic for u32). The
// kernel bodies are in scan_kernels.hpp, shared with the run-time compiled
// operator expressions.
//
// reduce: the reference's contract (detail::reduce, Algorithm.cpp:400-442):
// op(init, fold) with an order left unspecified (the operator must be
// commutative), n == 0 gives init, and the call blocks until the result is
// on the host. A grid of partial folds, then one workgroup folds the
// partials and applies init.
#include "context.hpp"
#include "ops.hpp"
#include "rtc.hpp"
#include "scan_kernels.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace clc {
namespace {

using namespace dev;

// ---- operators ---------------------------------------------------------------

template <typename T> struct wrap_type { using type = T; };
template <> struct wrap_type<int32_t> { using type = uint32_t; };
template <> struct wrap_type<int64_t> { using type = uint64_t; };

//! The canonical operators (reference Algorithm.hpp:53-65) plus the
//! projections. Integer plus/multiply wrap, as in OpenCL C.
template <typename T, int OP> struct canonical_op
{
    static __device__ __forceinline__ T apply(T a, T b)
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
};

template <typename T, int OP, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void scan_kernel(
    const T* in, T* out, size_t n, T identity, int exclusive, int vec_ok,
    uint64_t* status, uint32_t* tile_counter)
{
    scan_tile<T, canonical_op<T, OP>, BLOCK, ITEMS>(in, out, n, identity, exclusive,
                                                   vec_ok, status, tile_counter);
}

template <typename T, int OP, int BLOCK>
__global__ __launch_bounds__(BLOCK) void reduce_kernel(
    const T* __restrict__ in, size_t n, T identity, int vec_ok,
    T* __restrict__ out, int apply_init, T init)
{
    reduce_partial<T, canonical_op<T, OP>, BLOCK>(in, n, identity, vec_ok, out,
                                                 apply_init, init);
}

// ---- host side -------------------------------------------------------------------

//! Scratch bytes of a scan of n elements of T: the tile counter (256 B)
//! and one 16-byte look-back status slot per tile.
template <typename T> size_t scan_scratch_bytes(size_t n)
{
    constexpr size_t TILE = size_t(SCAN_BLOCK) * scan_items<T>();
    return 256 + (n + TILE - 1) / TILE * 16;
}

//! Clears the scratch t (scan_scratch_bytes<T>(n) bytes) with a kernel and
//! launches the scan on ctx's stream.
template <typename T, int OP>
void scan_launch(context& ctx, const void* in, void* out, size_t n,
                 const void* identity, bool exclusive, char* t)
{
    constexpr int BLOCK = SCAN_BLOCK;
    constexpr int ITEMS = scan_items<T>();
    constexpr size_t TILE = size_t(BLOCK) * ITEMS;
    const size_t tiles = (n + TILE - 1) / TILE;
    T id;
    std::memcpy(&id, identity, sizeof(T));
    const int vec_ok = (reinterpret_cast<uintptr_t>(in) % 16 == 0
                        && reinterpret_cast<uintptr_t>(out) % 16 == 0)
        ? 1 : 0;
    detail::zero_fill(ctx.stream(), t, scan_scratch_bytes<T>(n));
    hipLaunchKernelGGL((scan_kernel<T, OP, BLOCK, ITEMS>), dim3(tiles), dim3(BLOCK),
                       0, ctx.stream(), static_cast<const T*>(in),
                       static_cast<T*>(out), n, id, exclusive ? 1 : 0, vec_ok,
                       reinterpret_cast<uint64_t*>(t + 256),
                       reinterpret_cast<uint32_t*>(t));
    CLC_HIP(hipGetLastError());
}

template <typename T, int OP>
void scan_typed(context& ctx, const void* in, void* out, size_t n,
                const void* identity, bool exclusive)
{
    const size_t bytes = scan_scratch_bytes<T>(n);
    temp_buffer tmp(ctx.resource(), bytes, ctx.stream());
    timed_launch tl(ctx, "scan");
    scan_launch<T, OP>(ctx, in, out, n, identity, exclusive, tmp.as<char>());
}

template <typename T, int OP>
void reduce_typed(context& ctx, const void* in, size_t n, const void* identity,
                  const void* init, void* result)
{
    constexpr int BLOCK = REDUCE_BLOCK;
    T id, ini;
    std::memcpy(&id, identity, sizeof(T));
    std::memcpy(&ini, init, sizeof(T));
    const size_t per_block = size_t(BLOCK) * 16;
    const uint32_t groups = static_cast<uint32_t>(std::max<size_t>(
        1, std::min<size_t>((n + per_block - 1) / per_block,
                            size_t(ctx.facts().cu_count) * 8)));
    temp_buffer partials(ctx.resource(), (groups + 1) * sizeof(T), ctx.stream());
    T* part = partials.as<T>();
    const int vec_ok = reinterpret_cast<uintptr_t>(in) % 16 == 0 ? 1 : 0;
    {
        timed_launch tl(ctx, "reduce");
        hipLaunchKernelGGL((reduce_kernel<T, OP, BLOCK>), dim3(groups), dim3(BLOCK), 0,
                           ctx.stream(), static_cast<const T*>(in), n, id, vec_ok,
                           part, 0, ini);
    }
    {
        timed_launch tl(ctx, "reduce_partials");
        hipLaunchKernelGGL((reduce_kernel<T, OP, BLOCK>), dim3(1), dim3(BLOCK), 0,
                           ctx.stream(), part, size_t(groups), id, 0, part + groups,
                           1, ini);
    }
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

template <typename T, int OP> struct scan_fixed_fn
{
    static void run(context& ctx, const void* in, void* out, size_t n,
                    const void* id, bool ex, char* scratch_ptr, size_t capacity)
    {
        if (scan_scratch_bytes<T>(n) > capacity)
            throw std::length_error("clc_scan_noalloc: the call needs more than the "
                                    "fixed scratch holds");
        scan_launch<T, OP>(ctx, in, out, n, id, ex, scratch_ptr);
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

// ---- scan without allocation ------------------------------------------------
// Scratch of clc_scan_noalloc in a static device array, one copy per GPU
// (the loader places it with the code object; each GPU keeps its copy until
// the process ends). The largest call the C ABI takes, 2^31-1 elements of 8
// bytes at 8192 a tile, needs 256 + 262,144 x 16 bytes; smaller elements
// take fewer tiles, so every legal size fits. Host code reaches the array
// only through the address hipGetSymbolAddress returns on each device.
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 + (size_t(1) << 22);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mutex;
char* noalloc_scratch[NOALLOC_DEVICES] = {};

} // namespace

namespace detail {

hipError_t scan_noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (noalloc_scratch[device] != nullptr) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice; // the caller makes it current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        return e;
    }
    noalloc_scratch[device] = static_cast<char*>(p);
    return hipSuccess;
}

void scan_noalloc(context& ctx, const void* in, void* out, std::size_t n, int type,
                  int op, const void* identity, bool exclusive)
{
    if (n == 0) return;
    check_size(n);
    if (in == nullptr || out == nullptr)
        throw std::invalid_argument("clc_scan_noalloc: null buffer");
    ctx.activate();
    char* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(noalloc_mutex);
        if (ctx.device() >= 0 && ctx.device() < NOALLOC_DEVICES)
            p = noalloc_scratch[ctx.device()];
    }
    if (p == nullptr)
    {
        hip_check(scan_noalloc_setup(ctx.device()),
                  "clc_scan_noalloc: hipGetSymbolAddress of the scratch array");
        std::lock_guard<std::mutex> lock(noalloc_mutex);
        p = noalloc_scratch[ctx.device()];
    }
    by_type<scan_fixed_fn>(type, op, ctx, in, out, n, identity, exclusive, p,
                           NOALLOC_SCRATCH_BYTES);
}

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

void scan_expr(context& ctx, const void* in, void* out, std::size_t n, int type,
               const char* op, const char* identity, bool exclusive)
{
    const std::size_t es = type_size(type);
    if (op == nullptr || identity == nullptr)
        throw std::invalid_argument("clc_scan: null operator or identity");
    if (n == 0) return;
    check_size(n);
    if (in == nullptr || out == nullptr)
        throw std::invalid_argument("clc_scan: null buffer");
    ctx.activate();
    hipFunction_t f = ctx.programs().function(
        rtc::operator_source(type, op, identity, rtc::WITH_SCAN), rtc::SCAN_KERNEL);
    // Same tiling as the ahead-of-time scan.
    const std::size_t tile =
        std::size_t(SCAN_BLOCK) * scan_items_of(static_cast<unsigned>(es));
    const std::size_t tiles = (n + tile - 1) / tile;
    int vec_ok = (reinterpret_cast<uintptr_t>(in) % 16 == 0
                  && reinterpret_cast<uintptr_t>(out) % 16 == 0)
        ? 1 : 0;
    int ex = exclusive ? 1 : 0;
    const std::size_t bytes = 256 + tiles * 16;
    temp_buffer tmp(ctx.resource(), bytes, ctx.stream());
    char* t = tmp.as<char>();
    detail::zero_fill(ctx.stream(), t, bytes);
    void* status = t + 256;
    void* counter = t;
    void* args[] = { &in, &out, &n, &ex, &vec_ok, &status, &counter };
    timed_launch tl(ctx, "rtc_scan");
    CLC_HIP(hipModuleLaunchKernel(f, static_cast<unsigned>(tiles), 1, 1, SCAN_BLOCK, 1, 1,
                                  0, ctx.stream(), args, nullptr));
}

void reduce_expr(context& ctx, const void* in, std::size_t n, int type,
                 const char* op, const char* identity, const void* init,
                 void* result)
{
    const std::size_t es = type_size(type);
    if (op == nullptr || identity == nullptr)
        throw std::invalid_argument("clc_reduce: null operator or identity");
    if (n == 0)
    {
        std::memcpy(result, init, es);
        return;
    }
    check_size(n);
    if (in == nullptr) throw std::invalid_argument("clc_reduce: null buffer");
    ctx.activate();
    hipFunction_t f = ctx.programs().function(
        rtc::operator_source(type, op, identity, rtc::WITH_REDUCE), rtc::REDUCE_KERNEL);
    const std::size_t per_block = std::size_t(REDUCE_BLOCK) * 16;
    const uint32_t groups = static_cast<uint32_t>(std::max<std::size_t>(
        1, std::min<std::size_t>((n + per_block - 1) / per_block,
                                 std::size_t(ctx.facts().cu_count) * 8)));
    temp_buffer partials(ctx.resource(), (groups + 1) * es, ctx.stream());
    char* part = partials.as<char>();
    unsigned char ini[8] = {};
    std::memcpy(ini, init, es);
    {
        int vec_ok = reinterpret_cast<uintptr_t>(in) % 16 == 0 ? 1 : 0;
        int apply = 0;
        void* out = part;
        void* args[] = { &in, &n, &vec_ok, &out, &apply, ini };
        timed_launch tl(ctx, "rtc_reduce");
        CLC_HIP(hipModuleLaunchKernel(f, groups, 1, 1, REDUCE_BLOCK, 1, 1, 0,
                                      ctx.stream(), args, nullptr));
    }
    {
        const void* src = part;
        std::size_t count = groups;
        int vec_ok = 0, apply = 1;
        void* out = part + std::size_t(groups) * es;
        void* args[] = { &src, &count, &vec_ok, &out, &apply, ini };
        timed_launch tl(ctx, "rtc_reduce_partials");
        CLC_HIP(hipModuleLaunchKernel(f, 1, 1, 1, REDUCE_BLOCK, 1, 1, 0, ctx.stream(),
                                      args, nullptr));
    }
    CLC_HIP(hipMemcpyAsync(ctx.pinned(), part + std::size_t(groups) * es, es,
                           hipMemcpyDeviceToHost, ctx.stream()));
    CLC_HIP(hipStreamSynchronize(ctx.stream()));
    std::memcpy(result, ctx.pinned(), es);
}

std::size_t copy_if(context& ctx, const void* in, void* out, std::size_t n, int type,
                    const char* predicate)
{
    const std::size_t es = type_size(type);
    if (predicate == nullptr) throw std::invalid_argument("clc_copy_if: null predicate");
    if (n == 0) return 0;
    check_size(n);
    if (in == nullptr || out == nullptr)
        throw std::invalid_argument("clc_copy_if: null buffer");
    ctx.activate();
    hipFunction_t f = ctx.programs().function(rtc::predicate_source(type, predicate),
                                              rtc::COPY_IF_KERNEL);
    const std::size_t tile =
        std::size_t(SCAN_BLOCK) * scan_items_of(static_cast<unsigned>(es));
    const std::size_t tiles = (n + tile - 1) / tile;
    int vec_ok = reinterpret_cast<uintptr_t>(in) % 16 == 0 ? 1 : 0;
    const std::size_t bytes = 256 + tiles * 16;
    temp_buffer tmp(ctx.resource(), bytes, ctx.stream());
    char* t = tmp.as<char>();
    detail::zero_fill(ctx.stream(), t, bytes);
    void* status = t + 256;
    void* counter = t;
    void* count = t + 64;
    void* args[] = { &in, &out, &n, &vec_ok, &status, &counter, &count };
    {
        timed_launch tl(ctx, "rtc_copy_if");
        CLC_HIP(hipModuleLaunchKernel(f, static_cast<unsigned>(tiles), 1, 1, SCAN_BLOCK,
                                      1, 1, 0, ctx.stream(), args, nullptr));
    }
    CLC_HIP(hipMemcpyAsync(ctx.pinned(), count, 4, hipMemcpyDeviceToHost, ctx.stream()));
    CLC_HIP(hipStreamSynchronize(ctx.stream()));
    uint32_t kept = 0;
    std::memcpy(&kept, ctx.pinned(), 4);
    return kept;
}

} // namespace detail
} // namespace clc
