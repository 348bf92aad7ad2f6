// Kernels and launch sequence of the segmented sort (see segmented_sort.hpp
// for the design). Included by the per-key-width instantiation files.
#pragma once

#include "segmented_sort.hpp"

#include <algorithm>
#include <map>
#include <mutex>

namespace clc {
namespace seg {
namespace {

using namespace dev;
using namespace radix;

//! Bucket of a segment of m >= 2 keys: ceil(log2 m), capped.
__device__ __forceinline__ int bucket_of(uint64_t m)
{
    const int b = 64 - __clzll(m - 1);
    return b < NBUCKET ? b : NBUCKET - 1;
}

//! Files every segment of >= 2 keys under its class: an LDS class gets the
//! entry (begin low, begin high, size, digit bits) in its list (stride nseg), the
//! large class the segment's index. Segments that break the contract (end
//! < begin, end > num_items, 2^30 keys or more) are left untouched, and the
//! large class never takes more than large_cap segments, so malformed
//! offsets cannot drive any kernel out of bounds. The size limit is the
//! large path's: its look-back status words carry 30-bit digit counts.
__global__ __launch_bounds__(256) void classify_kernel(
    const uint64_t* __restrict__ off, uint32_t nseg, uint64_t num_items, table t,
    int use_mid, uint32_t large_cap, uint32_t* __restrict__ counts,
    uint4* __restrict__ entries, uint32_t* __restrict__ large_list)
{
    // The loop bound is uniform per block so that whole waves take part in
    // the per-class ballots: one atomic per class present in a wave (tens of
    // thousands of segments on a handful of counters would serialise).
    for (uint32_t base = blockIdx.x * 256; base < nseg; base += gridDim.x * 256)
    {
        const uint32_t i = base + threadIdx.x;
        int c = -1;
        uint32_t nb = MAX_DIGIT_BITS;
        uint64_t b = 0, e = 0;
        if (i < nseg)
        {
            b = off[i];
            e = off[i + 1];
            if (e > b + 1 && e <= num_items && e - b < (uint64_t(1) << 30))
            {
                const int bk = bucket_of(e - b);
                c = t.cls[bk];
                nb = static_cast<uint32_t>(t.digit_bits[bk]);
                if (use_mid && bk == MID_BUCKET && c < NCFG && e - b <= MID_CAP) c = MID;
            }
        }
        for (int k = 0; k <= MID; ++k)
        {
            const uint64_t mask = __ballot(c == k);
            if (mask == 0) continue;
            const uint32_t below = count_below(mask);
            uint32_t first = 0;
            if (c == k && below == 0)
                first = atomicAdd(&counts[k * CTR_STRIDE], uint32_t(__popcll(mask)));
            first = __shfl(first, __ffsll(static_cast<unsigned long long>(mask)) - 1, WAVE);
            if (c != k) continue;
            const uint32_t pos = first + below;
            if (k == LARGE)
            {
                if (pos < large_cap) large_list[pos] = i;
            }
            else // LDS classes; MID's list follows configuration NCFG - 1's
                entries[size_t(k == MID ? NCFG : k) * nseg + pos] =
                    make_uint4(static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32),
                               static_cast<uint32_t>(e - b), nb);
        }
    }
}

//! Persistent: workgroup w sorts the entries w, w + grid, w + 2 grid, ... of
//! its class (from classify_kernel), each wholly in LDS. The segments of a
//! class are within a factor of two in size, so the fixed stride balances
//! the work; a shared queue instead cost one device-scope atomic per
//! segment on one address, and those serialise (about 20,000 claims for
//! the smallest class of the 65,536 lognormal segments took longer than
//! the sorting). The entry is read by every thread from the same address
//! (a scalar load), so no LDS broadcast and no barrier is needed for it.
template <typename K, int KIND, bool DESC, bool VALUES, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void lds_class_kernel(
    K* keys, uint32_t* vals, const uint4* __restrict__ entries, const uint32_t* count)
{
    constexpr int W = BLOCK / WAVE;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem);
    void* s_buf = smem + W * 1024; // TILE * lds_segment_bytes_per_key bytes
    __shared__ uint32_t s_tmp[4];
    const uint32_t cnt = *count;
    for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x)
    {
        const uint4 e = entries[i];
        const uint64_t b = e.x | (static_cast<uint64_t>(e.y) << 32);
        const uint32_t nb = min(max(e.w, 1u), static_cast<uint32_t>(MAX_DIGIT_BITS));
        lds_sort_segment<K, KIND, DESC, VALUES, BLOCK, ITEMS>(
            keys + b, VALUES ? vals + b : nullptr, e.z, nb, s_hist, s_buf, s_tmp,
            i == blockIdx.x);
        __syncthreads(); // the next segment reuses the LDS
    }
}

//! Single block: lays out the onesweep tiles of the large segments.
//! Also clamps the large class to large_cap segments and max_tiles tiles
//! (only malformed offsets reach either bound) and publishes the number of
//! large segments in use.
__global__ __launch_bounds__(1024) void tile_scan_kernel(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ large_list,
    const uint32_t* large_count, uint32_t large_cap, uint32_t max_tiles, uint32_t tile,
    uint32_t* __restrict__ tile_start, uint32_t* large_used, uint32_t* total_tiles)
{
    __shared__ uint32_t s_wave[16];
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint32_t L = min(*large_count, large_cap);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < L; base += 1024)
    {
        const uint32_t i = base + tid;
        uint32_t nt = 0;
        if (i < L)
        {
            const uint32_t s = large_list[i];
            const uint64_t m = off[s + 1] - off[s];
            nt = static_cast<uint32_t>((m + tile - 1) / tile);
        }
        const uint32_t inc = wave_inclusive_add(nt);
        if (lane == WAVE - 1) s_wave[wave] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < 16; ++w)
        {
            if (w < static_cast<int>(wave)) before += s_wave[w];
            total += s_wave[w];
        }
        if (i < L) tile_start[i] = carry + before + inc - nt;
        carry += total;
        __syncthreads();
    }
    if (tid == 0)
    {
        tile_start[L] = carry;
        *large_used = L;
        // Tiles past max_tiles are dropped; a segment cut short this way
        // only ever scatters within itself.
        *total_tiles = min(carry, max_tiles);
    }
}

//! Tile layout -> segment map: map[t] = j for the tiles t of segment j
//! (start[j] <= t < start[j + 1], t < *total), so a tile finds its segment
//! with one load. (A binary search is about log2(count) dependent loads,
//! a few microseconds that a workgroup holding one tile per CU spends
//! idle on every tile.)
__global__ __launch_bounds__(256) void tile_map_kernel(const uint32_t* __restrict__ start,
                                                       const uint32_t* count,
                                                       const uint32_t* total,
                                                       uint32_t* __restrict__ map)
{
    const uint32_t L = *count, T = *total;
    for (uint32_t j = blockIdx.x; j < L; j += gridDim.x)
    {
        const uint32_t b = start[j], e = min(start[j + 1], T);
        for (uint32_t t = b + threadIdx.x; t < e; t += 256) map[t] = j;
    }
}

//! Persistent: per large segment, counts of every digit place. Workgroup
//! w takes tiles w, w + grid, ... (no shared queue: its device-scope
//! atomics on one address serialise, and the order does not matter here),
//! keeps counting in LDS while its tiles stay in one segment and adds the
//! counts to global memory when the segment changes and at the end (not
//! once per tile: every tile of a big segment would add to the same
//! P * 256 words). Each thread has U key loads in flight.
template <typename K, int KIND, bool DESC>
__global__ __launch_bounds__(256) void large_hist_kernel(
    const K* __restrict__ keys, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ large_list, const uint32_t* __restrict__ tile_start,
    const uint32_t* __restrict__ tile_seg, const uint32_t* total_tiles,
    uint32_t* __restrict__ hist)
{
    constexpr int P = sizeof(K);
    constexpr uint32_t TILE = large_tile(sizeof(K));
    constexpr int C = 32 / P; // interleaved copies per bin, as in the plain histogram
    constexpr uint32_t U = 8;
    __shared__ uint32_t s_h[P * RADIX * C];
    const uint32_t T = *total_tiles;
    const uint32_t tid = threadIdx.x, copy = tid % C;
    uint32_t cur = 0xFFFFFFFFu, first = 0;
    uint64_t b = 0, m = 0;
    auto flush = [&]() {
        for (uint32_t i = tid; i < P * RADIX; i += 256)
        {
            uint32_t c = 0;
#pragma unroll
            for (int q = 0; q < C; ++q) c += s_h[i * C + q];
            if (c != 0) atomicAdd(&hist[size_t(cur) * P * RADIX + i], c);
        }
    };
    for (uint32_t t = blockIdx.x; t < T; t += gridDim.x)
    {
        const uint32_t j = tile_seg[t]; // uniform
        if (j != cur)
        {
            __syncthreads(); // the previous segment's counts are complete
            if (cur != 0xFFFFFFFFu) flush();
            __syncthreads();
            for (uint32_t i = tid; i < P * RADIX * C; i += 256) s_h[i] = 0;
            cur = j;
            const uint32_t sidx = large_list[j];
            b = off[sidx];
            m = off[sidx + 1] - b;
            first = tile_start[j];
            __syncthreads();
        }
        const uint64_t lo = uint64_t(t - first) * TILE;
        const uint64_t hi = std::min<uint64_t>(lo + TILE, m);
        // Digits equal to the tile's first key's (a place constant across
        // the segment, common for large segments) are counted in registers
        // and added once: no LDS atomics all on one address.
        const K k0 = keys[b + lo];
        uint32_t dom[P], hits[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            dom[p] = digit_of<K, KIND, DESC>(k0, 8u * p);
            hits[p] = 0;
        }
        for (uint64_t i0 = lo; i0 < hi; i0 += 256 * U)
        {
            K k[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                k[u] = keys[b + std::min<uint64_t>(i0 + u * 256 + tid, hi - 1)]; // unconditional
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
            {
                if (i0 + u * 256 + tid >= hi) break;
#pragma unroll
                for (int p = 0; p < P; ++p)
                {
                    const uint32_t d = digit_of<K, KIND, DESC>(k[u], 8u * p);
                    if (d == dom[p]) ++hits[p];
                    else atomicAdd(&s_h[(p * RADIX + d) * C + copy], 1u);
                }
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p)
        {
            const uint32_t h = wave_sum(hits[p]);
            if (tid % WAVE == 0 && h != 0) atomicAdd(&s_h[(p * RADIX + dom[p]) * C], h);
        }
    }
    __syncthreads();
    if (cur != 0xFFFFFFFFu) flush();
}

//! Persistent: exclusive scan of each (large segment, digit place) count
//! row, and whether that place must run for that segment: a place where one
//! digit holds every key of the segment is skipped (identity permutation).
//! any_exec[p] (zeroed beforehand) is set when some segment runs place p.
__global__ __launch_bounds__(256) void large_digit_scan_kernel(
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ base,
    const uint32_t* large_used, uint32_t P, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ large_list, uint32_t* __restrict__ exec,
    uint32_t* __restrict__ any_exec)
{
    __shared__ uint32_t tmp[4];
    __shared__ uint32_t s_trivial;
    const uint32_t rows = *large_used * P;
    for (uint32_t r = blockIdx.x; r < rows; r += gridDim.x)
    {
        if (threadIdx.x == 0) s_trivial = 0;
        const size_t i = size_t(r) * RADIX + threadIdx.x;
        const uint32_t c = hist[i];
        base[i] = block256_exclusive_add<256>(c, tmp); // ends with a barrier
        const uint32_t s = large_list[r / P];
        if (uint64_t(c) == off[s + 1] - off[s]) s_trivial = 1;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            exec[r] = s_trivial ? 0u : 1u;
            if (!s_trivial) atomicOr(&any_exec[r % P], 1u);
        }
        __syncthreads();
    }
}

//! Single block: the large segments that run at least one digit place (the
//! others hold one repeated key and are already in order), as indices into
//! the large list, with their tile layout: act_start[a] = first tile of
//! active segment a, act_start[A] = all their tiles. The onesweep passes and
//! the copy-back walk only these tiles.
__global__ __launch_bounds__(1024) void active_scan_kernel(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ large_list,
    const uint32_t* large_used, const uint32_t* __restrict__ exec, uint32_t P,
    uint32_t max_tiles, uint32_t tile, uint32_t* __restrict__ active,
    uint32_t* __restrict__ act_start, uint32_t* act_used, uint32_t* act_tiles)
{
    __shared__ uint32_t s_f[16], s_t[16];
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const uint32_t L = *large_used;
    uint32_t carry_f = 0, carry_t = 0;
    for (uint32_t base = 0; base < L; base += 1024)
    {
        const uint32_t j = base + tid;
        uint32_t f = 0, nt = 0;
        if (j < L)
        {
            for (uint32_t p = 0; p < P; ++p) f |= exec[j * P + p];
            if (f)
            {
                const uint32_t s = large_list[j];
                nt = static_cast<uint32_t>((off[s + 1] - off[s] + tile - 1) / tile);
            }
        }
        const uint32_t inc_f = wave_inclusive_add(f), inc_t = wave_inclusive_add(nt);
        if (lane == WAVE - 1)
        {
            s_f[wave] = inc_f;
            s_t[wave] = inc_t;
        }
        __syncthreads();
        uint32_t bf = 0, bt = 0, tf = 0, tt = 0;
        for (int w = 0; w < 16; ++w)
        {
            if (w < static_cast<int>(wave))
            {
                bf += s_f[w];
                bt += s_t[w];
            }
            tf += s_f[w];
            tt += s_t[w];
        }
        if (f)
        {
            const uint32_t a = carry_f + bf + inc_f - 1;
            active[a] = j;
            act_start[a] = carry_t + bt + inc_t - nt;
        }
        carry_f += tf;
        carry_t += tt;
        __syncthreads();
    }
    if (tid == 0)
    {
        act_start[carry_f] = carry_t;
        *act_used = carry_f;
        *act_tiles = min(carry_t, max_tiles);
    }
}

//! Persistent over the large tiles: a segment that ran an odd number of
//! digit places holds its result in the scratch buffers; copy it back.
template <typename K, bool VALUES>
__global__ __launch_bounds__(256) void large_copy_back_kernel(
    K* __restrict__ keys, const K* __restrict__ keys_alt, uint32_t* __restrict__ vals,
    const uint32_t* __restrict__ vals_alt, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ large_list, const uint32_t* __restrict__ active,
    const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_seg,
    const uint32_t* total_tiles, const uint32_t* __restrict__ exec, uint32_t P)
{
    constexpr uint32_t TILE = large_tile(sizeof(K));
    const uint32_t T = *total_tiles;
    for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) // no queue, as large_hist_kernel
    {
        const uint32_t j = tile_seg[t];
        const uint32_t jf = active[j];
        uint32_t odd = 0;
        for (uint32_t p = 0; p < P; ++p) odd ^= exec[jf * P + p];
        if (!odd) continue;
        const uint32_t sidx = large_list[jf];
        const uint64_t b = off[sidx];
        const uint64_t m = off[sidx + 1] - b;
        const uint64_t lo = uint64_t(t - tile_start[j]) * TILE;
        const uint64_t hi = std::min<uint64_t>(lo + TILE, m);
        for (uint64_t i = lo + threadIdx.x; i < hi; i += 256)
        {
            keys[b + i] = keys_alt[b + i];
            if constexpr (VALUES) vals[b + i] = vals_alt[b + i];
        }
    }
}

//! Persistent onesweep over the tiles of every active large segment (tile
//! layout and index j from active_scan_kernel; jf = active[j] indexes the
//! large list), one digit place; look-back chains run within a segment.
//! Per segment, places whose digit is constant are skipped (exec[j][p] ==
//! 0) and each executed place reads what the previous executed place wrote
//! (a = keys, b = scratch). A place that no segment runs only clears the
//! next place's status words. Persistent rather than one workgroup per
//! tile: the tile count is only known on the device, and a grid sized for
//! the worst case would spend more on empty workgroups than on the tiles.
template <typename K, int KIND, bool DESC, bool VALUES>
__global__ __launch_bounds__(LARGE_BLOCK) void large_onesweep_kernel(
    K* keys_a, K* keys_b, uint32_t* vals_a, uint32_t* vals_b,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ large_list,
    const uint32_t* __restrict__ active, const uint32_t* __restrict__ tile_start,
    const uint32_t* __restrict__ tile_seg, const uint32_t* total_tiles, uint32_t pass,
    const uint32_t* __restrict__ base, const uint32_t* __restrict__ exec,
    const uint32_t* __restrict__ any_exec,
    uint32_t* status, uint32_t* status_next, uint32_t* queue)
{
    constexpr int P = sizeof(K);
    constexpr int W = LARGE_BLOCK / WAVE;
    constexpr int ITEMS = large_items(sizeof(K));
    constexpr uint32_t TILE = large_tile(sizeof(K));
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem);
    uint32_t* s_gbase = s_hist + W * RADIX;
    K* s_keys = reinterpret_cast<K*>(s_gbase + RADIX);
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys); // split staging
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_item, s_j;
    const uint32_t tid = threadIdx.x;
    const uint32_t T = *total_tiles;
    if (any_exec[pass] == 0)
    {
        if (status_next != nullptr)
            for (size_t i = size_t(blockIdx.x) * LARGE_BLOCK + tid; i < size_t(T) * RADIX;
                 i += size_t(gridDim.x) * LARGE_BLOCK)
                status_next[i] = 0;
        return;
    }
#pragma unroll 1
    for (;;)
    {
        if (tid == 0)
        {
            const uint32_t t = atomicAdd(queue, 1u);
            s_item = t;
            if (t < T) s_j = tile_seg[t];
        }
        for (uint32_t i = tid; i < W * RADIX; i += LARGE_BLOCK) s_hist[i] = 0;
        __syncthreads();
        const uint32_t t = s_item;
        if (t >= T) break;
        const uint32_t j = s_j, jf = active[j];
        if (status_next != nullptr && tid < RADIX) status_next[size_t(t) * RADIX + tid] = 0;
        if (exec[jf * P + pass] == 0)
        {
            __syncthreads(); // everyone has read s_item / s_j
            continue;
        }
        uint32_t parity = 0;
        for (uint32_t p = 0; p < pass; ++p) parity ^= exec[jf * P + p];
        const K* keys_in = parity ? keys_b : keys_a;
        K* keys_out = parity ? keys_a : keys_b;
        const uint32_t* vals_in = parity ? vals_b : vals_a;
        uint32_t* vals_out = parity ? vals_a : vals_b;
        const uint32_t sidx = large_list[jf];
        const uint64_t b = off[sidx];
        const uint32_t m = static_cast<uint32_t>(off[sidx + 1] - b);
        const uint32_t local = t - tile_start[j];
        const uint32_t* dbase = base + (size_t(jf) * P + pass) * RADIX;
        uint32_t* st = status + size_t(tile_start[j]) * RADIX;
        const uint32_t* vi = VALUES ? vals_in + b : nullptr;
        uint32_t* vo = VALUES ? vals_out + b : nullptr;
        const key_digit<K, KIND, DESC> dig{ 8u * pass };
        if (local * TILE + TILE <= m)
            onesweep_tile<K, uint32_t, VALUES, LARGE_BLOCK, ITEMS, true, true>(
                keys_in + b, keys_out + b, vi, vo, m, dig, dbase, st, local, s_hist, s_keys,
                s_vals, s_gbase, s_tmp);
        else
            onesweep_tile<K, uint32_t, VALUES, LARGE_BLOCK, ITEMS, false, true>(
                keys_in + b, keys_out + b, vi, vo, m, dig, dbase, st, local, s_hist, s_keys,
                s_vals, s_gbase, s_tmp);
        __syncthreads();
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

//! Resident workgroups of a kernel on the whole device (persistent grid),
//! cached per kernel.
template <typename F>
uint32_t resident_grid(const context& ctx, F kernel, int block, std::size_t lds)
{
    static std::mutex m;
    static std::map<const void*, uint32_t> cache;
    std::lock_guard<std::mutex> lock(m);
    const void* key = reinterpret_cast<const void*>(kernel);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds)
            != hipSuccess
        || per_cu < 1)
    {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const uint32_t g =
        static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(ctx.facts().cu_count);
    cache[key] = g;
    return g;
}

//! Persistent grid of an LDS shape (also opts the kernel in to more than
//! 64 KiB of LDS). Computed once per instantiation, outside any stream
//! capture.
template <typename K, int KIND, bool DESC, bool VALUES, int BLOCK, int ITEMS>
uint32_t lds_shape_grid(const context& ctx)
{
    static uint32_t grid = 0;
    static std::mutex m;
    std::lock_guard<std::mutex> lock(m);
    if (grid == 0)
    {
        auto kernel = lds_class_kernel<K, KIND, DESC, VALUES, BLOCK, ITEMS>;
        const std::size_t lds = shape_lds(BLOCK, std::size_t(BLOCK) * ITEMS, sizeof(K), VALUES);
        if (lds > 65536)
            CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        grid = resident_grid(ctx, kernel, BLOCK, lds);
    }
    return grid;
}

template <typename K, int KIND, bool DESC, bool VALUES, int BLOCK, int ITEMS>
void launch_lds_shape(context& ctx, bool prepare, K* keys, uint32_t* vals,
                      const uint4* entries, const uint32_t* count)
{
    const uint32_t grid = lds_shape_grid<K, KIND, DESC, VALUES, BLOCK, ITEMS>(ctx);
    if (prepare) return;
    static const std::string name =
        "segmented_lds_sort_" + std::to_string(BLOCK) + "x" + std::to_string(ITEMS);
    timed_launch timer(ctx, name.c_str());
    hipLaunchKernelGGL((lds_class_kernel<K, KIND, DESC, VALUES, BLOCK, ITEMS>), dim3(grid),
                       dim3(BLOCK), shape_lds(BLOCK, std::size_t(BLOCK) * ITEMS, sizeof(K), VALUES),
                       ctx.stream(), keys, vals, entries, count);
}

template <typename K, int KIND, bool DESC, bool VALUES, int C>
void launch_lds_class(context& ctx, bool prepare, K* keys, uint32_t* vals,
                      const uint4* entries, const uint32_t* count)
{
    launch_lds_shape<K, KIND, DESC, VALUES, CFG_BLOCK[C], CFG_ITEMS[C]>(
        ctx, prepare, keys, vals, entries, count);
}

template <typename K, int KIND, bool DESC, bool VALUES>
void launch_lds(context& ctx, bool prepare, int c, K* keys, uint32_t* vals,
                const uint4* entries, const uint32_t* count)
{
#define CLC_LDS_CASE(C)                                                        \
    case C:                                                                    \
        launch_lds_class<K, KIND, DESC, VALUES, C>(ctx, prepare, keys, vals,   \
                                                   entries, count);            \
        return;
    switch (c)
    {
        CLC_LDS_CASE(0)
        CLC_LDS_CASE(1)
        CLC_LDS_CASE(2)
        CLC_LDS_CASE(3)
        CLC_LDS_CASE(4)
        CLC_LDS_CASE(5)
        CLC_LDS_CASE(6)
    }
#undef CLC_LDS_CASE
    throw std::invalid_argument("segmented sort: configuration not available for this key type");
}

template <typename F> uint32_t cached_grid(const context& ctx, F kernel, int block)
{
    return resident_grid(ctx, kernel, block, 0);
}

//! Opts the large path's onesweep kernel in to more than 64 KiB of LDS,
//! once per kernel, outside any stream capture.
template <typename F> void large_lds_opt_in(F kernel, std::size_t lds)
{
    static std::mutex m;
    static std::map<const void*, bool> opted; // per kernel (one type, many kernels)
    {
        std::lock_guard<std::mutex> lock(m);
        const void* key = reinterpret_cast<const void*>(kernel);
        if (!opted[key] && lds > 65536)
            CLC_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        opted[key] = true;
    }
}

//! Enqueues the whole segmented sort on ctx's stream. With prepare set it
//! only computes the scratch size and the persistent grids (so a later
//! stream capture needs no runtime queries) and returns the size.
template <typename K, int KIND, bool DESC, bool VALUES>
std::size_t sort_impl(context& ctx, bool prepare, K* keys, uint32_t* vals,
                      std::size_t num_items, const uint64_t* off, std::size_t nseg,
                      const table& t, bool overlap_large)
{
    constexpr int P = sizeof(K);
    hipStream_t s = ctx.stream();
    int first_large = NBUCKET;
    bool used[NCFG] = {};
    for (int b = 1; b < NBUCKET; ++b)
    {
        if (t.cls[b] == LARGE) first_large = std::min(first_large, b);
        else if (t.cls[b] >= 0 && t.cls[b] < NCFG) used[t.cls[b]] = true;
        else throw std::invalid_argument("segmented sort: bad size-class table");
    }
    const bool use_mid = t.cls[MID_BUCKET] < NCFG
        && mid_lds(sizeof(K), VALUES) <= ctx.facts().lds_per_block;
    const std::size_t m_min_large =
        first_large < NBUCKET ? (std::size_t(1) << (first_large - 1)) + 1 : 0;
    const std::size_t L_max =
        first_large < NBUCKET ? std::min(nseg, num_items / m_min_large) : 0;
    constexpr uint32_t TILE = large_tile(sizeof(K));
    const std::size_t T_max = L_max ? (num_items + TILE - 1) / TILE + L_max : 0;

    layout Lt;
    const std::size_t o_ctr = Lt.take(CTR_SLOTS * CTR_STRIDE * 4);
    const std::size_t o_stA = Lt.take(T_max * RADIX * 4);
    const std::size_t o_hist = Lt.take(L_max * P * RADIX * 4);
    const std::size_t zero_bytes = Lt.bytes;
    const std::size_t o_entries = Lt.take(std::size_t(NCFG + 1) * nseg * 16);
    const std::size_t o_llist = Lt.take(nseg * 4);
    const std::size_t o_tstart = Lt.take((L_max + 1) * 4);
    const std::size_t o_active = Lt.take(L_max * 4);
    const std::size_t o_astart = Lt.take((L_max + 1) * 4);
    const std::size_t o_base = Lt.take(L_max * P * RADIX * 4);
    const std::size_t o_exec = Lt.take(L_max * P * 4);
    const std::size_t o_stB = Lt.take(T_max * RADIX * 4);
    const std::size_t o_tmap = Lt.take(T_max * 4);
    const std::size_t o_amap = Lt.take(T_max * 4);
    const std::size_t o_kalt = L_max ? Lt.take(num_items * sizeof(K)) : 0;
    const std::size_t o_valt = L_max && VALUES ? Lt.take(num_items * 4) : 0;

    auto hk = large_hist_kernel<K, KIND, DESC>;
    auto ok = large_onesweep_kernel<K, KIND, DESC, VALUES>;
    auto ck = large_copy_back_kernel<K, VALUES>;
    const uint32_t hgrid = L_max ? cached_grid(ctx, hk, 256) : 0;
    const std::size_t olds = large_lds(sizeof(K));
    if (L_max) large_lds_opt_in(ok, olds);
    const uint32_t ogrid = L_max ? resident_grid(ctx, ok, LARGE_BLOCK, olds) : 0;
    const uint32_t cpgrid = L_max ? cached_grid(ctx, ck, 256) : 0;
    if (prepare)
    {
        for (int c = 0; c < NCFG; ++c)
            if (used[c])
                launch_lds<K, KIND, DESC, VALUES>(ctx, true, c, nullptr, nullptr, nullptr,
                                                  nullptr);
        if (use_mid)
            launch_lds_shape<K, KIND, DESC, VALUES, MID_BLOCK, MID_ITEMS>(
                ctx, true, nullptr, nullptr, nullptr, nullptr);
        // The branch streams and events exist before any capture.
        for (int i = 0; i < context::AUX_STREAMS; ++i) ctx.aux_stream(i);
        for (int i = 0; i < context::AUX_EVENTS; ++i) ctx.aux_event(i);
        return Lt.bytes;
    }

    temp_buffer tmp(ctx.resource(), Lt.bytes, ctx.stream());
    char* base_ptr = tmp.as<char>();
    uint32_t* ctr = reinterpret_cast<uint32_t*>(base_ptr + o_ctr);
    // Counters, one slot of CTR_STRIDE words each (device-scope atomics on
    // words that share a slot serialise): slots [0, MID] list counts
    // (LARGE's at LARGE, MID's at MID), 17 tiles in use, 18 large segments
    // in use, 20 active large segments, 21 their tiles, [24, 24 + P)
    // onesweep queues (tiles must start in order there); slot 40 holds
    // the P flags of the places that some large segment runs, side by side
    // (set once each).
    static_assert(MID < 16 && 24 + 8 <= 40 && 40 < CTR_SLOTS, "counter layout");
    auto slot = [&](int i) { return ctr + std::size_t(i) * CTR_STRIDE; };
    uint32_t* counts = slot(0); // class k's count at counts[k * CTR_STRIDE]
    uint32_t* total_tiles = slot(17);
    uint32_t* large_used = slot(18);
    uint32_t* act_used = slot(20);
    uint32_t* act_tiles = slot(21);
    uint32_t* any_exec = slot(40);
    uint32_t* exec = reinterpret_cast<uint32_t*>(base_ptr + o_exec);
    uint4* entries = reinterpret_cast<uint4*>(base_ptr + o_entries);
    uint32_t* llist = reinterpret_cast<uint32_t*>(base_ptr + o_llist);
    uint32_t* tstart = reinterpret_cast<uint32_t*>(base_ptr + o_tstart);
    uint32_t* active = reinterpret_cast<uint32_t*>(base_ptr + o_active);
    uint32_t* astart = reinterpret_cast<uint32_t*>(base_ptr + o_astart);
    uint32_t* tmap = reinterpret_cast<uint32_t*>(base_ptr + o_tmap);
    uint32_t* amap = reinterpret_cast<uint32_t*>(base_ptr + o_amap);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base_ptr + o_hist);
    uint32_t* dbase = reinterpret_cast<uint32_t*>(base_ptr + o_base);
    uint32_t* st[2] = { reinterpret_cast<uint32_t*>(base_ptr + o_stA),
                        reinterpret_cast<uint32_t*>(base_ptr + o_stB) };

    detail::zero_fill(s, base_ptr, zero_bytes);
    const uint32_t cgrid = static_cast<uint32_t>(
        std::max<std::size_t>(1, std::min<std::size_t>((nseg + 255) / 256, 1024)));
    hipLaunchKernelGGL(classify_kernel, dim3(cgrid), dim3(256), 0, s, off,
                       static_cast<uint32_t>(nseg), static_cast<uint64_t>(num_items), t,
                       use_mid ? 1 : 0, static_cast<uint32_t>(L_max), counts, entries, llist);

    // The large path (segments above every LDS capacity), on stream ls.
    auto large_path = [&](hipStream_t ls) {
        hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, ls, off, llist,
                           slot(LARGE), static_cast<uint32_t>(L_max),
                           static_cast<uint32_t>(T_max), TILE, tstart, large_used, total_tiles);
        const uint32_t mgrid = static_cast<uint32_t>(std::min<std::size_t>(L_max, 2048));
        hipLaunchKernelGGL(tile_map_kernel, dim3(mgrid), dim3(256), 0, ls, tstart, large_used,
                           total_tiles, tmap);
        hipLaunchKernelGGL(hk, dim3(hgrid), dim3(256), 0, ls, keys, off, llist, tstart, tmap,
                           total_tiles, hist);
        const uint32_t sgrid = static_cast<uint32_t>(std::min<std::size_t>(L_max * P, 4096));
        hipLaunchKernelGGL(large_digit_scan_kernel, dim3(sgrid), dim3(256), 0, ls, hist,
                           dbase, large_used, uint32_t(P), off, llist, exec, any_exec);
        hipLaunchKernelGGL(active_scan_kernel, dim3(1), dim3(1024), 0, ls, off, llist,
                           large_used, exec, uint32_t(P), static_cast<uint32_t>(T_max), TILE,
                           active, astart, act_used, act_tiles);
        hipLaunchKernelGGL(tile_map_kernel, dim3(mgrid), dim3(256), 0, ls, astart, act_used,
                           act_tiles, amap);
        K* kb[2] = { keys, reinterpret_cast<K*>(base_ptr + o_kalt) };
        uint32_t* vb[2] = { vals, VALUES ? reinterpret_cast<uint32_t*>(base_ptr + o_valt)
                                         : nullptr };
        for (int p = 0; p < P; ++p)
        {
            timed_launch timer(ctx, "segmented_onesweep");
            hipLaunchKernelGGL(ok, dim3(ogrid), dim3(LARGE_BLOCK), olds, ls, kb[0], kb[1],
                               vb[0], vb[1], off, llist, active, astart, amap,
                               act_tiles, uint32_t(p), dbase, exec, any_exec, st[p % 2],
                               p + 1 < P ? st[(p + 1) % 2] : nullptr, slot(24 + p));
        }
        hipLaunchKernelGGL(ck, dim3(cpgrid), dim3(256), 0, ls, kb[0], kb[1], vb[0], vb[1],
                           off, llist, active, astart, amap, act_tiles, exec,
                           uint32_t(P));
    };

    // The LDS classes sort disjoint segments: each is a branch forked from
    // stream s onto an auxiliary stream and joined back, so the classes'
    // persistent kernels overlap (a workgroup whose share of its class is
    // empty just exits) instead of running one after another with a
    // drained GPU between them. Captured into a graph, the branches become
    // parallel nodes. The classes fork from the largest down, the mid
    // class last (forked first it measured slower: 7.04 against 6.61 ms for
    // the learned sort of 2^28 lognormal keys, whose buckets fill the mid
    // class and three others; the classes then compete for the CUs from
    // the start). The large path runs on s: with overlap_large at once,
    // beside the classes (the branches are joined after it), otherwise
    // once they are done (see sort_typed for the measurements of both).
    hipEvent_t fork = ctx.aux_event(0);
    CLC_HIP(hipEventRecord(fork, s));
    int branch = 0;
    struct stream_swap
    {
        context& c;
        hipStream_t prev;
        ~stream_swap() { c.swap_stream(prev); }
    };
    hipEvent_t pending[context::AUX_EVENTS];
    int npending = 0;
    auto run_branch = [&](auto&& body) {
        const hipStream_t as = ctx.aux_stream(branch % context::AUX_STREAMS);
        CLC_HIP(hipStreamWaitEvent(as, fork, 0));
        {
            stream_swap guard{ ctx, ctx.swap_stream(as) };
            body(as);
        }
        const hipEvent_t done = ctx.aux_event(1 + branch);
        CLC_HIP(hipEventRecord(done, as));
        if (overlap_large) pending[npending++] = done;
        else CLC_HIP(hipStreamWaitEvent(s, done, 0));
        ++branch;
    };
    static_assert(NCFG + 2 < context::AUX_EVENTS, "one event per branch and the fork");
    for (int c = NCFG - 1; c >= 0; --c)
        if (used[c])
            run_branch([&](hipStream_t) {
                launch_lds<K, KIND, DESC, VALUES>(ctx, false, c, keys, vals,
                                                  entries + std::size_t(c) * nseg, slot(c));
            });
    if (use_mid)
        run_branch([&](hipStream_t) {
            launch_lds_shape<K, KIND, DESC, VALUES, MID_BLOCK, MID_ITEMS>(
                ctx, false, keys, vals, entries + std::size_t(NCFG) * nseg, slot(MID));
        });
    if (L_max > 0) large_path(s);
    for (int i = 0; i < npending; ++i) CLC_HIP(hipStreamWaitEvent(s, pending[i], 0));
    CLC_HIP(hipGetLastError());
    return Lt.bytes;
}

} // namespace

template <typename K, int KIND>
std::size_t sort_typed(context& ctx, bool prepare, void* keys, bool with_values,
                       void* values, std::size_t num_items, const uint64_t* offsets,
                       std::size_t num_segments, bool descending, const table& t,
                       bool overlap_large)
{
    K* k = static_cast<K*>(keys);
    uint32_t* v = static_cast<uint32_t*>(values);
    const std::size_t n = num_items, ns = num_segments;
    if (descending)
    {
        if (with_values) return sort_impl<K, KIND, true, true>(ctx, prepare, k, v, n, offsets, ns, t, overlap_large);
        return sort_impl<K, KIND, true, false>(ctx, prepare, k, v, n, offsets, ns, t, overlap_large);
    }
    if (with_values) return sort_impl<K, KIND, false, true>(ctx, prepare, k, v, n, offsets, ns, t, overlap_large);
    return sort_impl<K, KIND, false, false>(ctx, prepare, k, v, n, offsets, ns, t, overlap_large);
}

} // namespace seg
} // namespace clc
