// Stable LSD radix sort for gfx950: 8-bit digits, one upfront histogram of
// every digit place, then one onesweep kernel per digit place.
//
// Same contract as the reference's detail::radix_sort
// (lib/src/Compute/Algorithm.cpp:453-527): in place, stable, keys ordered by
// the KEY_TRANSFORM image (:459-469, here order_bits), values permuted
// alongside. The algorithm is different: the reference runs 8 passes of
// 4-bit digits, each a histogram launch, a recursive scan and a scatter
// whose ranks come from a work-group scan of packed counters
// (Algorithm.cpp:129-213). Here:
//
//   digit_histogram  one read of the keys counts all sizeof(K) digit places
//                    at once (LDS counters, one global add per bin per block)
//   digit_scan       exclusive scan of each place's 256 counts
//   onesweep (x P)   per tile: ranks from per-wave digit counters in LDS
//                    (an atomic add per key, or a ballot match for rounds of
//                    mostly equal digits), decoupled look-back across tiles per
//                    digit, tile reordered in LDS by digit and written to HBM
//                    in contiguous digit runs.
//
// Tiles take their index from an atomic counter when they start, so a tile
// only ever waits on tiles that are already running (no co-residency
// assumption). Ranks follow input order inside a wave (lane order within a
// 64-key round, rounds in order), across waves (wave-ordered prefixes) and
// across tiles (look-back), which keeps the sort stable. A digit place where
// every key has the same digit is skipped on the device (its pass would be
// the identity); each executed pass reads what the previous one wrote. The
// result ends in the caller's buffers (in place, as Algorithm.cpp:526): an
// odd number of executed passes routes through a second scratch buffer so
// the last pass writes there (radix::pass_route), and only a single executed
// pass, or a pool without room for that buffer, needs a final copy. No host
// round trip is needed.
#include "context.hpp"
#include "device_common.hpp"
#include "ops.hpp"
#include "radix_onesweep.hpp"
#include "radix_tile.hpp"
#include "scan_kernels.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace clc {
namespace {

using namespace dev;
using namespace radix;

// ---- upfront histogram of every digit place -------------------------------

//! Every bin has C = 32 KiB / (P KiB) copies, interleaved (copy = lane mod
//! C, adjacent copies in adjacent banks): lanes that hit the same digit,
//! the common case for skewed keys, serialise at most 64 / C deep instead
//! of 64. 64-bit keys (whose high places are often constant: keys below
//! 2^48, lognormal or Zipf keys) also count the digit of keys[0] at each
//! place in a register (a constant place's digit; no LDS atomic for it)
//! and add it once at the end.
template <typename K, int KIND, bool DESC, int BLOCK, bool VEC>
__global__ __launch_bounds__(BLOCK) void digit_histogram_kernel(
    const K* __restrict__ keys, uint32_t n, uint32_t* __restrict__ hist)
{
    constexpr int P = sizeof(K);
    constexpr int C = 32 / P;
    constexpr bool DOM = P == 8;
    __shared__ uint32_t s_hist[P * RADIX * C];
    for (uint32_t i = threadIdx.x; i < P * RADIX * C; i += BLOCK) s_hist[i] = 0;
    const uint32_t copy = threadIdx.x % C;
    uint32_t dom[DOM ? P : 1], hits[DOM ? P : 1];
    if constexpr (DOM)
    {
        const K k0 = keys[0];
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            dom[p] = digit_of<K, KIND, DESC>(k0, 8u * p);
            hits[p] = 0;
        }
    }
    __syncthreads();
    auto count = [&](K k) {
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            const uint32_t d = digit_of<K, KIND, DESC>(k, 8u * p);
            if constexpr (DOM)
                if (d == dom[p])
                {
                    ++hits[p];
                    continue;
                }
            atomicAdd(&s_hist[(p * RADIX + d) * C + copy], 1u);
        }
    };
    const uint32_t stride = gridDim.x * BLOCK;
    uint32_t start = 0;
    if constexpr (VEC)
    {
        constexpr uint32_t PER = 16 / sizeof(K);
        const uint32_t nvec = n / PER;
        const uint4* v = reinterpret_cast<const uint4*>(keys);
        // Two 16-byte loads per thread per round, and the next round's two
        // issued before this round's LDS atomics (which do not wait on
        // memory), so four are in flight while the atomics run. Loads past
        // the end are clamped to the last vector and not counted.
        uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
        if (nvec != 0 && i < nvec)
        {
            uint4 q0 = load_stream(v + i);
            uint4 q1 = load_stream(v + min(i + stride, nvec - 1));
            for (;;)
            {
                const uint32_t next = i + 2 * stride;
                const uint4 r0 = load_stream(v + min(next, nvec - 1));
                const uint4 r1 = load_stream(v + min(next + stride, nvec - 1));
                const K* k0 = reinterpret_cast<const K*>(&q0);
                const K* k1 = reinterpret_cast<const K*>(&q1);
#pragma unroll
                for (uint32_t j = 0; j < PER; ++j) count(k0[j]);
                if (i + stride < nvec)
#pragma unroll
                    for (uint32_t j = 0; j < PER; ++j) count(k1[j]);
                if (next >= nvec) break;
                i = next;
                q0 = r0;
                q1 = r1;
            }
        }
        start = nvec * PER;
    }
    for (uint32_t i = start + blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride)
        count(keys[i]);
    if constexpr (DOM)
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            const uint32_t h = wave_sum(hits[p]);
            if (threadIdx.x % WAVE == 0 && h != 0)
                atomicAdd(&s_hist[(p * RADIX + dom[p]) * C], h);
        }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < P * RADIX; i += BLOCK)
    {
        uint32_t c = 0;
#pragma unroll
        for (int j = 0; j < C; ++j) c += s_hist[i * C + j];
        if (c != 0) atomicAdd(&hist[i], c);
    }
}

//! base[p][d] = sum of hist[p][d'] for d' < d; one block per digit place.
//! A place where one digit holds all n keys is marked trivial (its pass
//! would be the identity permutation; trivial[] starts zeroed).
__global__ __launch_bounds__(256) void digit_scan_kernel(
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ base, uint32_t n,
    uint32_t* __restrict__ trivial)
{
    __shared__ uint32_t tmp[4];
    const uint32_t i = blockIdx.x * RADIX + threadIdx.x;
    const uint32_t c = hist[i];
    base[i] = block256_exclusive_add<256>(c, tmp);
    if (c == n) trivial[blockIdx.x] = 1;
}

//! After the passes: where the routing of the passes left the result in the
//! scratch buffers b (an odd number of executed passes, and either just one
//! or no second scratch c; see radix::pass_route), copy it back to the
//! caller's a.
template <typename K>
__global__ __launch_bounds__(256) void finish_copy_kernel(
    K* __restrict__ keys_a, const K* __restrict__ keys_b, uint32_t* __restrict__ vals_a,
    const uint32_t* __restrict__ vals_b, uint32_t n, const uint32_t* __restrict__ trivial,
    uint32_t passes, uint32_t three)
{
    uint32_t ran = 0;
    for (uint32_t p = 0; p < passes; ++p) ran += trivial[p] == 0 ? 1u : 0u;
    if ((ran & 1u) == 0 || (three && ran > 1)) return;
    const uint32_t stride = gridDim.x * 256;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    {
        keys_a[i] = keys_b[i];
        if (vals_a != nullptr) vals_a[i] = vals_b[i];
    }
}

struct layout
{
    std::size_t bytes = 0;
    std::size_t take(std::size_t b)
    {
        const std::size_t o = bytes;
        bytes += (b + 255) & ~std::size_t(255);
        return o;
    }
};

//! Tile shape for a key width and values flag: the CLC_ONESWEEP_CFG
//! environment variable, else the tuned shape, else shape 0; a shape whose
//! LDS does not fit this key width is never used.
int pick_shape(context& ctx, std::size_t kbytes, bool values)
{
    if (const char* e = std::getenv("CLC_ONESWEEP_CFG"))
    {
        const int c = std::atoi(e);
        if (os_feasible(c, kbytes, values)) return c;
    }
    auto& tuned = ctx.tuned();
    auto it = tuned.find(detail::sort_tuning_slot(kbytes, values));
    if (it != tuned.end() && it->second.size() == 1
        && os_feasible(it->second[0], kbytes, values))
        return it->second[0];
    std::vector<int> stored;
    if (load_tuning(ctx, detail::sort_tuning_name(kbytes, values), stored)
        && stored.size() == 1 && os_feasible(stored[0], kbytes, values))
    {
        tuned[detail::sort_tuning_slot(kbytes, values)] = stored;
        return stored[0];
    }
    return os_default(kbytes, values);
}

template <typename K, int KIND>
void sort_impl(context& ctx, K* keys, uint32_t* vals, uint32_t n, bool desc, int shape)
{
    constexpr int P = sizeof(K);
    const bool values = vals != nullptr;
    const bool wide = n >= (1u << 30); // 30-bit counts fit a 32-bit status word
    const std::size_t sbytes = wide ? 8 : 4;
    const uint32_t tiles = (n + os_tile(shape) - 1) / os_tile(shape);
    hipStream_t s = ctx.stream();

    layout L;
    const std::size_t o_hist = L.take(P * RADIX * 4);
    const std::size_t o_ctr = L.take(P * 4);
    const std::size_t o_triv = L.take(P * 4);
    const std::size_t o_st0 = L.take(std::size_t(tiles) * RADIX * sbytes);
    const std::size_t zero_bytes = L.bytes;
    const std::size_t o_st1 = L.take(std::size_t(tiles) * RADIX * sbytes);
    const std::size_t o_base = L.take(P * RADIX * 4);
    const std::size_t o_kalt = L.take(std::size_t(n) * sizeof(K));
    const std::size_t o_valt = values ? L.take(std::size_t(n) * 4) : 0;
    // A second scratch pair (c) lets an odd number of executed passes end in
    // the caller's buffers (radix::pass_route); where the pool cannot give
    // the memory, the passes alternate between two buffers and an odd count
    // is copied back at the end (about two thirds of a pass).
    layout L3 = L;
    const std::size_t o_kalt2 = L3.take(std::size_t(n) * sizeof(K));
    const std::size_t o_valt2 = values ? L3.take(std::size_t(n) * 4) : 0;
    std::unique_ptr<temp_buffer> tmp;
    bool three = true;
    try
    {
        tmp = std::make_unique<temp_buffer>(ctx.resource(), L3.bytes, ctx.stream());
    }
    catch (const out_of_memory&)
    {
        three = false;
        tmp = std::make_unique<temp_buffer>(ctx.resource(), L.bytes, ctx.stream());
    }
    char* t = tmp->as<char>();
    uint32_t* hist = reinterpret_cast<uint32_t*>(t + o_hist);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(t + o_ctr);
    uint32_t* trivial = reinterpret_cast<uint32_t*>(t + o_triv);
    void* st[2] = { t + o_st0, t + o_st1 };
    uint32_t* base = reinterpret_cast<uint32_t*>(t + o_base);
    K* kb[3] = { keys, reinterpret_cast<K*>(t + o_kalt),
                 three ? reinterpret_cast<K*>(t + o_kalt2) : nullptr };
    uint32_t* vb[3] = { vals, values ? reinterpret_cast<uint32_t*>(t + o_valt) : nullptr,
                        values && three ? reinterpret_cast<uint32_t*>(t + o_valt2) : nullptr };

    detail::zero_fill(s, t, zero_bytes);
    const uint32_t hist_grid =
        std::max(1u, std::min<uint32_t>((n + 4095) / 4096,
                                        uint32_t(ctx.facts().cu_count) * 8));
    {
        timed_launch timer(ctx, "radix_digit_histogram");
        const bool vec = reinterpret_cast<uintptr_t>(keys) % 16 == 0;
#define CLC_HIST(D, V)                                                          \
    hipLaunchKernelGGL((digit_histogram_kernel<K, KIND, D, 256, V>), dim3(hist_grid), \
                       dim3(256), 0, s, keys, n, hist)
        if (desc) { if (vec) CLC_HIST(true, true); else CLC_HIST(true, false); }
        else { if (vec) CLC_HIST(false, true); else CLC_HIST(false, false); }
#undef CLC_HIST
    }
    hipLaunchKernelGGL(digit_scan_kernel, dim3(P), dim3(256), 0, s, hist, base, n,
                       trivial);
    for (int p = 0; p < P; ++p)
    {
        timed_launch timer(ctx, values ? "radix_onesweep_pairs" : "radix_onesweep_keys");
        onesweep_args a{ kb[0], kb[1], kb[2], vb[0], vb[1], vb[2], n, uint32_t(p), trivial,
                         base + p * RADIX, st[p % 2],
                         p + 1 < P ? st[(p + 1) % 2] : nullptr, ctr + p,
                         uint32_t(ctx.facts().cu_count) };
        launch_onesweep<K, KIND>(shape, desc, values, wide, a, s);
    }
    const uint32_t copy_grid =
        std::max(1u, std::min<uint32_t>((n + 255) / 256, uint32_t(ctx.facts().cu_count) * 4));
    hipLaunchKernelGGL((finish_copy_kernel<K>), dim3(copy_grid), dim3(256), 0, s, kb[0],
                       kb[1], vb[0], vb[1], n, trivial, uint32_t(P), three ? 1u : 0u);
    CLC_HIP(hipGetLastError());
}

using sort_fn = void (*)(context&, void*, void*, uint32_t, bool, int);

template <typename K, int KIND>
void sort_entry(context& ctx, void* keys, void* values, uint32_t n, bool desc, int shape)
{
    sort_impl<K, KIND>(ctx, static_cast<K*>(keys), static_cast<uint32_t*>(values), n,
                       desc, shape);
}

sort_fn pick_sort(int key_type)
{
    switch (key_type)
    {
        case TY_U32: return &sort_entry<uint32_t, KEY_UNSIGNED>;
        case TY_I32: return &sort_entry<uint32_t, KEY_SIGNED>;
        case TY_F32: return &sort_entry<uint32_t, KEY_FLOAT>;
        case TY_U64: return &sort_entry<uint64_t, KEY_UNSIGNED>;
        case TY_I64: return &sort_entry<uint64_t, KEY_SIGNED>;
        case TY_F64: return &sort_entry<uint64_t, KEY_FLOAT>;
    }
    throw std::invalid_argument("clc_sort: keys must be I32, U32, F32, I64, U64 or F64");
}

std::size_t key_bytes_of(int key_type)
{
    return (key_type == TY_U64 || key_type == TY_I64 || key_type == TY_F64) ? 8 : 4;
}

__global__ void fill_words_kernel(uint32_t* w, std::size_t n, uint64_t seed)
{
    for (std::size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += size_t(gridDim.x) * 256)
    {
        uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull; // splitmix64
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        w[i] = static_cast<uint32_t>(z ^ (z >> 31));
    }
}

} // namespace

namespace detail {

int sort_tuning_slot(std::size_t kbytes, bool values)
{
    return 2000 + static_cast<int>(kbytes) * 2 + (values ? 1 : 0);
}

std::string sort_tuning_name(std::size_t kbytes, bool values)
{
    return "sort_k" + std::to_string(kbytes * 8) + (values ? "_kv" : "_k");
}

void radix_sort(context& ctx, void* keys, int key_type, void* values,
                std::size_t n, bool descending)
{
    const sort_fn fn = pick_sort(key_type);
    if (n <= 1) return;
    check_size(n);
    if (keys == nullptr) throw std::invalid_argument("clc_sort: null keys");
    ctx.activate();
    const int shape = pick_shape(ctx, key_bytes_of(key_type), values != nullptr);
    fn(ctx, keys, values, static_cast<uint32_t>(n), descending, shape);
}

int sort_tune(context& ctx, int key_type, bool with_values)
{
    const sort_fn fn = pick_sort(key_type);
    ctx.activate();
    const std::size_t kb = key_bytes_of(key_type);
    constexpr uint32_t N = 1u << 26; // 512 MiB of pairs: past the 256 MB MALL
    temp_buffer keys(ctx.resource(), std::size_t(N) * kb, ctx.stream());
    temp_buffer vals(ctx.resource(), with_values ? std::size_t(N) * 4 : 0, ctx.stream());
    hipEvent_t e0, e1;
    CLC_HIP(hipEventCreate(&e0));
    CLC_HIP(hipEventCreate(&e1));
    struct guard
    {
        hipEvent_t a, b;
        ~guard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } g{ e0, e1 };
    hipStream_t s = ctx.stream();
    const bool profiling = ctx.profiler().enabled();
    ctx.profiler().enable(false);
    int best = 0;
    float best_ms = 1e30f;
    for (int shape = 0; shape < NOS; ++shape)
    {
        if (!os_feasible(shape, kb, with_values)) continue;
        float shape_ms = 1e30f;
        for (int rep = 0; rep < 4; ++rep)
        {
            hipLaunchKernelGGL(fill_words_kernel, dim3(2048), dim3(256), 0, s,
                               keys.as<uint32_t>(), std::size_t(N) * kb / 4,
                               uint64_t(rep + 1));
            CLC_HIP(hipEventRecord(e0, s));
            fn(ctx, keys.get(), with_values ? vals.get() : nullptr, N, false, shape);
            CLC_HIP(hipEventRecord(e1, s));
            CLC_HIP(hipEventSynchronize(e1));
            float ms = 0;
            CLC_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) shape_ms = std::min(shape_ms, ms); // rep 0 warms up
        }
        if (shape_ms < best_ms)
        {
            best_ms = shape_ms;
            best = shape;
        }
    }
    ctx.profiler().enable(profiling);
    ctx.tuned()[sort_tuning_slot(kb, with_values)] = { best };
    store_tuning(ctx, sort_tuning_name(kb, with_values), { best });
    return best;
}

} // namespace detail
} // namespace clc
