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
//   onesweep (x P)   per tile: wave64 ballot match-and-rank, per-wave digit
//                    counters in LDS, decoupled look-back across tiles per
//                    digit, tile reordered in LDS by digit and written to HBM
//                    in contiguous digit runs.
//
// Tiles take their index from an atomic counter when they start, so a tile
// only ever waits on tiles that are already running (no co-residency
// assumption). Ranks follow input order inside a wave (lane order within a
// 64-key round, rounds in order), across waves (wave-ordered prefixes) and
// across tiles (look-back), which keeps the sort stable. P = sizeof(K)
// passes is even, so the result ends in the caller's buffers.
#include "context.hpp"
#include "device_common.hpp"
#include "radix_tile.hpp"

#include <algorithm>

namespace clc {
namespace {

using namespace dev;
using namespace radix;

// ---- upfront histogram of every digit place -------------------------------

template <typename K, int KIND, bool DESC, int BLOCK, bool VEC>
__global__ __launch_bounds__(BLOCK) void digit_histogram_kernel(
    const K* __restrict__ keys, uint32_t n, uint32_t* __restrict__ hist)
{
    constexpr int P = sizeof(K);
    __shared__ uint32_t s_hist[P * RADIX];
    for (uint32_t i = threadIdx.x; i < P * RADIX; i += BLOCK) s_hist[i] = 0;
    __syncthreads();
    auto count = [&](K k) {
#pragma unroll
        for (int p = 0; p < P; ++p)
            atomicAdd(&s_hist[p * RADIX + digit_of<K, KIND, DESC>(k, 8u * p)], 1u);
    };
    const uint32_t stride = gridDim.x * BLOCK;
    uint32_t start = 0;
    if constexpr (VEC)
    {
        constexpr uint32_t PER = 16 / sizeof(K);
        const uint32_t nvec = n / PER;
        const uint4* v = reinterpret_cast<const uint4*>(keys);
        for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < nvec; i += stride)
        {
            const uint4 q = v[i];
            const K* ks = reinterpret_cast<const K*>(&q);
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j) count(ks[j]);
        }
        start = nvec * PER;
    }
    for (uint32_t i = start + blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride)
        count(keys[i]);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < P * RADIX; i += BLOCK)
        if (s_hist[i] != 0) atomicAdd(&hist[i], s_hist[i]);
}

//! base[p][d] = sum of hist[p][d'] for d' < d; one block per digit place.
__global__ __launch_bounds__(256) void digit_scan_kernel(
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ base)
{
    __shared__ uint32_t tmp[4];
    const uint32_t i = blockIdx.x * RADIX + threadIdx.x;
    base[i] = block256_exclusive_add<256>(hist[i], tmp);
}

template <typename K, typename S, int KIND, bool DESC, bool VALUES, int BLOCK,
          int ITEMS>
__global__ __launch_bounds__(BLOCK) void onesweep_kernel(
    const K* __restrict__ keys_in, K* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out,
    uint32_t n, uint32_t shift, const uint32_t* __restrict__ digit_base,
    S* status, S* status_next, uint32_t* tile_counter)
{
    static_assert(BLOCK >= 256 && BLOCK % WAVE == 0, "block shape");
    constexpr int W = BLOCK / WAVE;
    constexpr uint32_t TILE = BLOCK * ITEMS;

    __shared__ uint32_t s_hist[W * RADIX]; // per-wave digit counters
    __shared__ K s_keys[TILE]; // tile reordered by digit
    __shared__ uint32_t s_vals[VALUES ? TILE : 1];
    __shared__ uint32_t s_gbase[RADIX]; // global destination - tile digit start
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    for (uint32_t i = tid; i < W * RADIX; i += BLOCK) s_hist[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    // Clear this tile's status words for the next digit place; that pass
    // starts after this one has finished, so nobody reads them meanwhile.
    if (status_next != nullptr && tid < RADIX)
        status_next[static_cast<size_t>(tile) * RADIX + tid] = 0;

    if (tile * TILE + TILE <= n)
        onesweep_tile<K, S, VALUES, BLOCK, ITEMS, true>(
            keys_in, keys_out, vals_in, vals_out, n, key_digit<K, KIND, DESC>{ shift }, digit_base, status,
            tile, s_hist, s_keys, s_vals, s_gbase, s_tmp);
    else
        onesweep_tile<K, S, VALUES, BLOCK, ITEMS, false>(
            keys_in, keys_out, vals_in, vals_out, n, key_digit<K, KIND, DESC>{ shift }, digit_base, status,
            tile, s_hist, s_keys, s_vals, s_gbase, s_tmp);
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

template <typename K, typename S, int KIND, bool DESC, bool VALUES>
void sort_impl(context& ctx, K* keys, uint32_t* vals, uint32_t n)
{
    constexpr int P = sizeof(K);
    constexpr int BLOCK = 256, ITEMS = 16;
    constexpr uint32_t TILE = BLOCK * ITEMS;
    const uint32_t tiles = (n + TILE - 1) / TILE;
    hipStream_t s = ctx.stream();

    layout L;
    const std::size_t o_hist = L.take(P * RADIX * 4);
    const std::size_t o_ctr = L.take(P * 4);
    const std::size_t o_st0 = L.take(std::size_t(tiles) * RADIX * sizeof(S));
    const std::size_t zero_bytes = L.bytes;
    const std::size_t o_st1 = L.take(std::size_t(tiles) * RADIX * sizeof(S));
    const std::size_t o_base = L.take(P * RADIX * 4);
    const std::size_t o_kalt = L.take(std::size_t(n) * sizeof(K));
    const std::size_t o_valt = VALUES ? L.take(std::size_t(n) * 4) : 0;
    temp_buffer tmp(ctx.resource(), L.bytes);
    char* t = tmp.as<char>();
    uint32_t* hist = reinterpret_cast<uint32_t*>(t + o_hist);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(t + o_ctr);
    S* st[2] = { reinterpret_cast<S*>(t + o_st0), reinterpret_cast<S*>(t + o_st1) };
    uint32_t* base = reinterpret_cast<uint32_t*>(t + o_base);
    K* kb[2] = { keys, reinterpret_cast<K*>(t + o_kalt) };
    uint32_t* vb[2] = { vals, VALUES ? reinterpret_cast<uint32_t*>(t + o_valt)
                                     : nullptr };

    detail::zero_fill(s, t, zero_bytes);
    const uint32_t hist_grid =
        std::max(1u, std::min<uint32_t>((n + 4095) / 4096,
                                        uint32_t(ctx.facts().cu_count) * 8));
    {
        timed_launch timer(ctx, "radix_digit_histogram");
        if (reinterpret_cast<uintptr_t>(keys) % 16 == 0)
            hipLaunchKernelGGL((digit_histogram_kernel<K, KIND, DESC, 256, true>),
                               dim3(hist_grid), dim3(256), 0, s, keys, n, hist);
        else
            hipLaunchKernelGGL((digit_histogram_kernel<K, KIND, DESC, 256, false>),
                               dim3(hist_grid), dim3(256), 0, s, keys, n, hist);
    }
    hipLaunchKernelGGL(digit_scan_kernel, dim3(P), dim3(256), 0, s, hist, base);
    for (int p = 0; p < P; ++p)
    {
        const int src = p % 2, dst = 1 - src;
        timed_launch timer(ctx, VALUES ? "radix_onesweep_pairs" : "radix_onesweep_keys");
        hipLaunchKernelGGL(
            (onesweep_kernel<K, S, KIND, DESC, VALUES, BLOCK, ITEMS>),
            dim3(tiles), dim3(BLOCK), 0, s, kb[src], kb[dst], vb[src], vb[dst],
            n, uint32_t(8 * p), base + p * RADIX, st[p % 2],
            p + 1 < P ? st[(p + 1) % 2] : nullptr, ctr + p);
    }
    CLC_HIP(hipGetLastError());
}

template <typename K, int KIND>
void sort_dispatch(context& ctx, void* keys, void* values, uint32_t n, bool desc)
{
    K* k = static_cast<K*>(keys);
    uint32_t* v = static_cast<uint32_t*>(values);
    const bool wide = n >= (1u << 30);
#define CLC_SORT_CASE(S, D, V)                                                 \
    sort_impl<K, S, KIND, D, V>(ctx, k, v, n)
    if (wide)
    {
        if (desc) { if (v) CLC_SORT_CASE(uint64_t, true, true); else CLC_SORT_CASE(uint64_t, true, false); }
        else { if (v) CLC_SORT_CASE(uint64_t, false, true); else CLC_SORT_CASE(uint64_t, false, false); }
    }
    else
    {
        if (desc) { if (v) CLC_SORT_CASE(uint32_t, true, true); else CLC_SORT_CASE(uint32_t, true, false); }
        else { if (v) CLC_SORT_CASE(uint32_t, false, true); else CLC_SORT_CASE(uint32_t, false, false); }
    }
#undef CLC_SORT_CASE
}

} // namespace

namespace detail {

void radix_sort(context& ctx, void* keys, int key_type, void* values,
                std::size_t n, bool descending)
{
    switch (key_type)
    {
        case 4: case 5: case 8: case 6: case 7: case 9: break;
        default:
            throw std::invalid_argument(
                "clc_sort: keys must be I32, U32, F32, I64, U64 or F64");
    }
    if (n <= 1) return;
    check_size(n);
    if (keys == nullptr) throw std::invalid_argument("clc_sort: null keys");
    ctx.activate();
    const uint32_t m = static_cast<uint32_t>(n);
    switch (key_type)
    {
        case 5: sort_dispatch<uint32_t, KEY_UNSIGNED>(ctx, keys, values, m, descending); break;
        case 4: sort_dispatch<uint32_t, KEY_SIGNED>(ctx, keys, values, m, descending); break;
        case 8: sort_dispatch<uint32_t, KEY_FLOAT>(ctx, keys, values, m, descending); break;
        case 7: sort_dispatch<uint64_t, KEY_UNSIGNED>(ctx, keys, values, m, descending); break;
        case 6: sort_dispatch<uint64_t, KEY_SIGNED>(ctx, keys, values, m, descending); break;
        case 9: sort_dispatch<uint64_t, KEY_FLOAT>(ctx, keys, values, m, descending); break;
    }
}

} // namespace detail
} // namespace clc
