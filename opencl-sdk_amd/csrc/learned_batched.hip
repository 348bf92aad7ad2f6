// Batched learned-bucket sort for gfx950: many segments, one learned CDF
// model per segment, every stage a fixed launch over all segments (no host
// round trip). No counterpart in the reference, whose only sort is
// detail::radix_sort (Algorithm.cpp:453-527); the per-segment equivalent is
// one plain sort per segment, and this library's segmented sort stays the
// fallback.
//
//   prepare      one block: per segment its size, whether it is learned
//                (more keys than the LDS segment sort holds, at least
//                2^15), its sample
//                count (1%), its bucket count (one per ~3072 keys) and its
//                slots in the bucket layout; exclusive scans of these and
//                the onesweep tile layout of the learned segments
//   sample       a strided sample of each learned segment's order images
//                (sample_count: every 100th key up to 65,536, every 400th beyond)
//   sort         the segmented sort over the samples (one segment each)
//   fit          one launch, one workgroup per segment: the segment's RMI
//                (root linear model over the order image to 256 leaves,
//                leaves fitted one per thread over the sorted sample), and
//                the fit check (sample keys in buckets estimated overfull
//                holding more than one distinct key); a segment that fails
//                it keeps one bucket (the segmented sort's onesweep path)
//   histogram    per learned segment, counts of the low and high byte of
//                the bucket id (inference fused in)
//   2 passes     onesweep passes over the learned segments' tiles keyed by
//                the low, then the high byte of the bucket id, look-back
//                chains within a segment; a segment with at most 256
//                buckets runs the first only
//   bounds       every bucket's first key by binary search, written as the
//                offsets of a new segment list: the buckets of learned
//                segments, every other segment whole
//   finish       the segmented sort over that list
//
// Model and monotonicity arguments are the single-array learned sort's
// (learned_model.hpp): equal keys share a bucket, bucket order is key
// order, and every pass is stable, so the result is the stable sort of each
// segment, bit for bit.
#include "context.hpp"
#include "device_common.hpp"
#include "learned_model.hpp"
#include "learned_sort.hpp"
#include "ops.hpp"
#include "radix_tile.hpp"
#include "segmented_sort.hpp"

#include <algorithm>
#include <cstring>

namespace clc {
namespace learned {
namespace {

using namespace dev;
using namespace radix;

//! The root's secant runs through the sample's quantiles 1/ROOT_QUANTILE
//! and 1 - 1/ROOT_QUANTILE.
constexpr uint32_t ROOT_QUANTILE = 512;
//! Target keys per bucket: buckets near it fit the 4096-key LDS class.
constexpr uint32_t BUCKET_TARGET = 3072;
//! Smallest segment given a model (327 samples); smaller ones, and those
//! the LDS segment sort holds, are sorted whole. Below 2^30 keys: 30-bit
//! look-back counts.
constexpr uint32_t MIN_LEARNED = 1u << 15;
constexpr uint32_t MAX_BUCKETS = 65536; // 16-bit bucket ids: two 8-bit passes
//! Onesweep tile of the id passes (split staging, digits staged in LDS): the
//! most keys a thread that fit 128 VGPRs without spilling (24 spill); 34.3-34.7
//! ms for the batched float segments against 35.1 ms with 18.
constexpr int BBLOCK = 1024, BITEMS = 22;
constexpr uint32_t BTILE = BBLOCK * BITEMS;

//! Per segment, filled by prepare_kernel.
struct seg_info
{
    unsigned long long begin; // first key
    uint32_t m; // keys (0 for a malformed segment)
    uint32_t learned; // 1: fitted (set by prepare), 0: sorted whole
    uint32_t nb; // buckets (learned segments)
    uint32_t slot; // first slot in the bucket layout
    uint32_t tile; // first onesweep tile (learned segments)
    uint32_t pad;
};

//! Single block: per segment, its bounds, whether it is learned (m > cap,
//! m >= MIN_LEARNED, m < 2^30),
//! its sample count and bucket count, and exclusive scans over segments of
//! samples (sample_off, u64 for the segmented sort), layout slots (a
//! learned segment has nb, any other one) and onesweep tiles. totals =
//! {samples, slots, tiles}.
__global__ __launch_bounds__(1024) void prepare_kernel(
    const uint64_t* __restrict__ off, uint32_t nseg, uint64_t num_items, uint32_t cap,
    seg_info* __restrict__ info, uint64_t* __restrict__ sample_off, uint32_t* __restrict__ totals)
{
    __shared__ uint32_t s_w[3][16];
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    uint32_t carry[3] = { 0, 0, 0 };
    for (uint32_t base = 0; base < nseg; base += 1024)
    {
        const uint32_t j = base + tid;
        uint32_t x[3] = { 0, 0, 0 }; // samples, slots, tiles
        seg_info si{};
        if (j < nseg)
        {
            const uint64_t b = off[j], e = off[j + 1];
            const bool ok = b <= e && e <= num_items && e - b < (1ull << 31);
            si.begin = b;
            si.m = ok ? static_cast<uint32_t>(e - b) : 0u;
            si.learned = si.m > cap && si.m >= MIN_LEARNED && si.m < (1u << 30) ? 1u : 0u;
            if (si.learned)
            {
                si.nb = min(MAX_BUCKETS, (si.m + BUCKET_TARGET - 1) / BUCKET_TARGET);
                x[0] = sample_count(si.m);
                x[1] = si.nb;
                x[2] = (si.m + BTILE - 1) / BTILE;
            }
            else
            {
                si.nb = 1;
                x[1] = 1;
            }
        }
        uint32_t inc[3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
        {
            inc[q] = wave_inclusive_add(x[q]);
            if (lane == WAVE - 1) s_w[q][wave] = inc[q];
        }
        __syncthreads();
        uint32_t before[3] = { 0, 0, 0 }, total[3] = { 0, 0, 0 };
        for (int w = 0; w < 16; ++w)
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                if (w < static_cast<int>(wave)) before[q] += s_w[q][w];
                total[q] += s_w[q][w];
            }
        if (j < nseg)
        {
            const uint32_t ex0 = carry[0] + before[0] + inc[0] - x[0];
            sample_off[j] = ex0;
            si.slot = carry[1] + before[1] + inc[1] - x[1];
            si.tile = carry[2] + before[2] + inc[2] - x[2];
            info[j] = si;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) carry[q] += total[q];
        __syncthreads();
    }
    if (tid == 0)
    {
        sample_off[nseg] = carry[0];
        totals[0] = carry[0];
        totals[1] = carry[1];
        totals[2] = carry[2];
    }
}

//! Block per segment: the strided sample of a learned segment (order
//! images; they sort ascending as unsigned 32-bit keys).
template <int KIND, bool DESC>
__global__ __launch_bounds__(256) void sample_kernel(const uint32_t* __restrict__ keys,
                                                     const seg_info* __restrict__ info,
                                                     const uint64_t* __restrict__ sample_off,
                                                     uint32_t* __restrict__ samples)
{
    const seg_info si = info[blockIdx.x];
    if (!si.learned) return;
    const uint32_t ns = sample_count(si.m), every = si.m / ns; // si.m >= MIN_LEARNED
    uint32_t* out = samples + sample_off[blockIdx.x];
    const uint32_t* in = keys + si.begin;
    for (uint32_t i = threadIdx.x; i < ns; i += 256)
        out[i] = order_bits<uint32_t, KIND, DESC>(in[size_t(i) * every + every / 2]);
}

//! Block per learned segment: fits its RMI on the sorted sample and checks
//! the fit. The root is linear in the 32-bit order image (for floats that
//! image is already piecewise linear in log2 |x|): the secant of the
//! sample's CDF between two quantiles. Each leaf is fitted by
//! one wave over its samples (least squares relative to the leaf's first
//! sample). The check counts samples i whose bucket also holds sample i + K,
//! K samples standing for cap keys, with a different key: those keys sit in
//! buckets too large for LDS with more than one distinct key. Above
//! fit_limit of the sample the segment is sorted whole (learned = 0).
__global__ __launch_bounds__(256) void fit_kernel(const uint32_t* __restrict__ samples,
                                                  const uint64_t* __restrict__ sample_off,
                                                  seg_info* __restrict__ info,
                                                  rmi_model* __restrict__ models, uint32_t cap,
                                                  double fit_limit,
                                                  unsigned long long* __restrict__ fit_bad)
{
    __shared__ uint32_t lstart[LEAVES + 1];
    __shared__ rmi_model sm;
    __shared__ uint32_t s_cnt[4];
    seg_info* sip = info + blockIdx.x;
    if (!sip->learned) return; // uniform per block
    const uint32_t m = sip->m, nb = sip->nb;
    const uint64_t so = sample_off[blockIdx.x];
    const uint32_t ns = static_cast<uint32_t>(sample_off[blockIdx.x + 1] - so);
    const uint32_t* s = samples + so;
    const uint32_t tid = threadIdx.x;

    // Root: the secant of the sample's CDF through its quantiles q and 1 - q
    // (q = 1/ROOT_QUANTILE), as in the single-array learned sort: the
    // samples between them spread over the leaves by rank.
    if (tid == 0)
    {
        const uint32_t i1 = ns / ROOT_QUANTILE, i2 = ns - 1 - i1;
        const double scale = double(LEAVES) / double(ns);
        if (s[i2] > s[i1])
        {
            sm.a0 = double(i2 - i1) * scale / double(s[i2] - s[i1]);
            sm.b0 = double(i1) * scale - sm.a0 * static_cast<double>(s[i1]);
        }
        else
        {
            sm.a0 = 0.0;
            sm.b0 = double(LEAVES / 2);
        }
        sm.nb = double(nb);
        sm.root_log = 0;
        sm.pad = 0;
    }
    __syncthreads();

    // Leaf ranges of the sorted sample.
    for (uint32_t i = tid; i < ns; i += 256)
    {
        const uint32_t li = leaf_of(sm, double(s[i]));
        const int lp = i == 0 ? -1 : int(leaf_of(sm, double(s[i - 1])));
        for (int l = lp + 1; l <= int(li); ++l) lstart[l] = i;
        if (i == ns - 1)
            for (uint32_t l = li + 1; l <= LEAVES; ++l) lstart[l] = ns;
    }
    __syncthreads();

    // Leaves, one wave each at a time (leaf wave, wave + 4, ...): the lanes
    // stride over the leaf's samples, the moments are summed over the wave
    // in a fixed order, lane 0 stores the leaf. (One thread per leaf walked
    // a large leaf's samples one dependent load after another.)
    const uint32_t lane = tid % WAVE;
    for (uint32_t l = tid / WAVE; l < LEAVES; l += 256 / WAVE)
    {
        const uint32_t lo = lstart[l], hi = lstart[l + 1];
        double a = 0, b = (double(lo) + double(hi)) / (2.0 * ns);
        if (hi - lo >= 2)
        {
            double lm[5] = { 0, 0, 0, 0, 0 };
            for (uint32_t i = lo + lane; i < hi; i += WAVE)
            {
                const double x = static_cast<double>(s[i] - s[lo]);
                const double y = static_cast<double>(i - lo);
                lm[0] += 1.0;
                lm[1] += x;
                lm[2] += y;
                lm[3] += x * x;
                lm[4] += x * y;
            }
#pragma unroll
            for (int k = 0; k < 5; ++k)
#pragma unroll
                for (int o = WAVE / 2; o > 0; o >>= 1) lm[k] += __shfl_xor(lm[k], o, WAVE);
            double ai, ci;
            line_from_moments(lm, ai, ci); // index - lo ~ ai * (x - s[lo]) + ci
            a = ai / double(ns);
            b = (double(lo) + ci - ai * static_cast<double>(s[lo])) / double(ns);
        }
        if (lane == 0)
        {
            const uint32_t blo = static_cast<uint32_t>(uint64_t(nb) * lo / ns);
            const uint32_t next =
                l + 1 < LEAVES ? static_cast<uint32_t>(uint64_t(nb) * hi / ns) : nb - 1;
            rmi_leaf& L = sm.leaf[l];
            L.a = a * double(nb); // positions in buckets, not CDF fractions
            L.b = b * double(nb);
            L.lo = min(blo, nb - 1);
            L.hi = max(L.lo, min(next, nb - 1));
        }
    }
    __syncthreads();

    // Fit check.
    const uint32_t K = max(1u, static_cast<uint32_t>(uint64_t(cap) * ns / m));
    uint32_t over = 0;
    for (uint32_t i = tid; i + K < ns; i += 256)
        if (s[i] != s[i + K] && bucket_of(sm, s[i]) == bucket_of(sm, s[i + K])) ++over;
    over = wave_sum(over);
    if (tid % WAVE == 0) s_cnt[tid / WAVE] = over;
    __syncthreads();
    const uint32_t total = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    const bool good = double(total) <= fit_limit * double(ns);

    // Publish: the model, or one bucket.
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&sm);
    uint32_t* dst = reinterpret_cast<uint32_t*>(models + blockIdx.x);
    for (uint32_t i = tid; i < sizeof(rmi_model) / 4; i += 256) dst[i] = src[i];
    if (tid == 0 && !good)
    {
        sip->learned = 0;
        atomicAdd(fit_bad, 1ull);
    }
}

//! Tile -> segment map of the id passes' tile layout: tile_seg[t] = j for
//! the tiles of each segment that prepare_kernel made learned (run right
//! after it, before a fit can clear the flag: the layout stays). One load
//! per tile instead of a binary search over the segments, about log2(nseg)
//! dependent loads spent idle by every tile.
__global__ __launch_bounds__(256) void tile_map_kernel(const seg_info* __restrict__ info,
                                                       uint32_t nseg,
                                                       uint32_t* __restrict__ tile_seg)
{
    for (uint32_t j = blockIdx.x; j < nseg; j += gridDim.x)
    {
        const seg_info si = info[j];
        if (!si.learned) continue;
        const uint32_t nt = (si.m + BTILE - 1) / BTILE;
        for (uint32_t k = threadIdx.x; k < nt; k += 256) tile_seg[si.tile + k] = j;
    }
}

//! Per learned segment (j), counts of the bucket id's low and high byte:
//! hist[j][0][d], hist[j][1][d]. Block b walks tiles [b * HU, b * HU + HU)
//! in order, counting in LDS and adding to global memory when the segment
//! changes or at the end (one flush per segment piece, not per tile), with
//! HL independent loads per thread in flight. Both bytes count into HC
//! interleaved copies per bin (copy = lane mod HC), so lanes that share a
//! bucket do not serialise on one LDS word (the high byte takes few values;
//! it is 0 for segments of at most 256 buckets, whose second pass does not
//! run, and is not counted there). One ballot per high-byte value and wave
//! instead took the batched sort of 4096 float segments (2.08e9 pairs) from
//! 40.1 to 42.2 ms: the kernel is bound by its instructions (the
//! inference), and the ballots cost more of them than the atomic.
constexpr uint32_t HU = 8, HL = 4, HC = 8;
template <int KIND, bool DESC>
__global__ __launch_bounds__(1024) void id_hist_kernel(const uint32_t* __restrict__ keys,
                                                       const seg_info* __restrict__ info,
                                                       const uint32_t* __restrict__ tile_seg,
                                                       const rmi_model* __restrict__ models,
                                                       const uint32_t* totals,
                                                       uint32_t* __restrict__ hist)
{
    __shared__ rmi_model sm;
    __shared__ uint32_t s_h[2 * RADIX * HC];
    const uint32_t T = totals[2];
    const uint32_t t0 = blockIdx.x * HU, t1 = min(T, t0 + HU);
    const uint32_t copy = threadIdx.x % HC;
    uint32_t cur = 0xFFFFFFFFu; // segment in LDS
    auto flush = [&](uint32_t j) {
        for (uint32_t i = threadIdx.x; i < 2 * RADIX; i += 1024)
        {
            uint32_t c = 0;
#pragma unroll
            for (uint32_t q = 0; q < HC; ++q) c += s_h[i * HC + q];
            if (c != 0) atomicAdd(&hist[size_t(j) * 2 * RADIX + i], c);
        }
    };
    for (uint32_t t = t0; t < t1; ++t)
    {
        const uint32_t j = tile_seg[t];
        const seg_info si = info[j];
        if (!si.learned) continue; // uniform: every thread reads the same info
        if (j != cur)
        {
            __syncthreads(); // previous counts complete
            if (cur != 0xFFFFFFFFu) flush(cur);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < 2 * RADIX * HC; i += 1024) s_h[i] = 0;
            load_model(&sm, models + j);
            cur = j;
            __syncthreads();
        }
        const uint32_t hmax = (si.nb - 1) >> 8; // high bytes lie in [0, hmax]
        const uint32_t lo = (t - si.tile) * BTILE, hi = min(si.m, lo + BTILE);
        const uint32_t* in = keys + si.begin;
        for (uint32_t i0 = lo; i0 < hi; i0 += HL * 1024)
        {
            const uint32_t i = i0 + threadIdx.x;
            uint32_t k[HL];
#pragma unroll
            for (uint32_t u = 0; u < HL; ++u) k[u] = in[min(i + u * 1024, hi - 1)];
#pragma unroll
            for (uint32_t u = 0; u < HL; ++u)
            {
                if (i + u * 1024 < hi)
                {
                    const uint32_t b = bucket_of(sm, order_bits<uint32_t, KIND, DESC>(k[u]));
                    atomicAdd(&s_h[(b & 0xFFu) * HC + copy], 1u);
                    if (hmax > 0) atomicAdd(&s_h[(RADIX + (b >> 8)) * HC + copy], 1u);
                }
            }
        }
    }
    __syncthreads();
    if (cur != 0xFFFFFFFFu) flush(cur);
}

//! Block per (segment, place): exclusive scan of the 256 counts.
__global__ __launch_bounds__(256) void id_scan_kernel(const uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ base)
{
    __shared__ uint32_t tmp[4];
    const size_t i = size_t(blockIdx.x) * RADIX + threadIdx.x;
    base[i] = block256_exclusive_add<256>(hist[i], tmp);
}

//! One id pass over every learned segment's tiles: pass 0 keyed by the
//! bucket id's low byte (keys -> alt), pass 1 by its high byte (alt -> keys,
//! segments with more than 256 buckets only). Tiles take their index from
//! an atomic counter (a tile waits only on tiles already running), the
//! look-back runs within the tile's segment.
template <int KIND, bool DESC, bool VALUES>
__global__ __launch_bounds__(BBLOCK) void id_onesweep_kernel(
    const uint32_t* __restrict__ keys_in, uint32_t* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out,
    const seg_info* __restrict__ info, const uint32_t* __restrict__ tile_seg,
    const rmi_model* __restrict__ models, const uint32_t* totals, uint32_t pass, const uint32_t* __restrict__ base,
    uint32_t* status, uint32_t* status_next, uint32_t* tile_counter)
{
    constexpr int W = BBLOCK / WAVE;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem);
    uint32_t* s_gbase = s_hist + W * RADIX;
    rmi_model* s_model = reinterpret_cast<rmi_model*>(s_gbase + RADIX);
    uint32_t* s_keys = reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(s_model)
                                                   + ((sizeof(rmi_model) + 15) & ~size_t(15)));
    uint8_t* s_dig = reinterpret_cast<uint8_t*>(s_keys + BTILE);
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= totals[2]) return; // uniform
    if (status_next != nullptr && tid < RADIX) status_next[size_t(t) * RADIX + tid] = 0;
    const uint32_t j = tile_seg[t];
    const seg_info si = info[j];
    if (!si.learned || (pass == 1 && si.nb <= RADIX)) return; // this pass leaves it as is
    for (uint32_t i = tid; i < W * RADIX; i += BBLOCK) s_hist[i] = 0;
    load_model(s_model, models + j);
    __syncthreads();
    const uint32_t local = t - si.tile;
    const id_digit<uint32_t, KIND, DESC> dig{ s_model, 8u * pass, s_dig };
    const uint32_t* dbase = base + (size_t(j) * 2 + pass) * RADIX;
    uint32_t* st = status + size_t(si.tile) * RADIX;
    const uint32_t* vi = VALUES ? vals_in + si.begin : nullptr;
    uint32_t* vo = VALUES ? vals_out + si.begin : nullptr;
    uint32_t* s_vals = s_keys; // split staging: values after keys
    if (local * BTILE + BTILE <= si.m)
        onesweep_tile<uint32_t, uint32_t, VALUES, BBLOCK, BITEMS, true, true>(
            keys_in + si.begin, keys_out + si.begin, vi, vo, si.m, dig, dbase, st, local, s_hist,
            s_keys, s_vals, s_gbase, s_tmp);
    else
        onesweep_tile<uint32_t, uint32_t, VALUES, BBLOCK, BITEMS, false, true>(
            keys_in + si.begin, keys_out + si.begin, vi, vo, si.m, dig, dbase, st, local, s_hist,
            s_keys, s_vals, s_gbase, s_tmp);
}

//! Block per tile: learned segments that ran only the first id pass hold
//! their keys in the scratch buffers; copy them back.
template <bool VALUES>
__global__ __launch_bounds__(256) void copy_back_kernel(
    uint32_t* __restrict__ keys, const uint32_t* __restrict__ keys_alt,
    uint32_t* __restrict__ vals, const uint32_t* __restrict__ vals_alt,
    const seg_info* __restrict__ info, const uint32_t* __restrict__ tile_seg,
    const uint32_t* totals)
{
    const uint32_t t = blockIdx.x;
    if (t >= totals[2]) return;
    const uint32_t j = tile_seg[t];
    const seg_info si = info[j];
    if (!si.learned || si.nb > RADIX) return;
    const uint32_t lo = (t - si.tile) * BTILE, hi = min(si.m, lo + BTILE);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256)
    {
        keys[si.begin + i] = keys_alt[si.begin + i];
        if constexpr (VALUES) vals[si.begin + i] = vals_alt[si.begin + i];
    }
}

//! Thread per layout slot: the offsets of the bucket list. Slot c of a
//! learned segment starts at its first key whose bucket is >= c (binary
//! search; keys are grouped by non-decreasing bucket); a segment sorted
//! whole has one slot at its begin; slots past the layout and the end
//! entry take end_off. Malformed segments keep their offsets, so the
//! segmented sort skips them as it would have.
template <int KIND, bool DESC>
__global__ __launch_bounds__(256) void bounds_kernel(const uint32_t* __restrict__ keys,
                                                     const seg_info* __restrict__ info,
                                                     uint32_t nseg,
                                                     const rmi_model* __restrict__ models,
                                                     const uint64_t* __restrict__ off,
                                                     const uint32_t* totals, uint32_t slots,
                                                     uint64_t* __restrict__ out)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q > slots) return;
    const uint32_t used = totals[1];
    if (q >= used)
    {
        out[q] = off[nseg];
        return;
    }
    uint32_t lo = 0, hi = nseg; // segment: last j with info[j].slot <= q
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) / 2;
        if (info[mid].slot <= q) lo = mid;
        else hi = mid;
    }
    const uint32_t j = lo;
    const seg_info si = info[j];
    const uint32_t c = q - si.slot;
    if (si.m == 0)
    {
        out[q] = off[j]; // malformed or empty: the original offset
        return;
    }
    if (!si.learned)
    {
        out[q] = c == 0 ? si.begin : si.begin + si.m;
        return;
    }
    const rmi_model& M = models[j];
    const uint32_t* in = keys + si.begin;
    uint32_t a = 0, b = si.m; // first index with bucket >= c, in [a, b]
    while (a < b)
    {
        const uint32_t mid = a + (b - a) / 2;
        if (bucket_of(M, order_bits<uint32_t, KIND, DESC>(in[mid])) >= c) b = mid;
        else a = mid + 1;
    }
    out[q] = si.begin + a;
}

//! stats[0] = keys in buckets (or whole segments) above cap, stats[1] =
//! largest bucket of a learned segment.
__global__ __launch_bounds__(256) void bucket_stats_kernel(const uint64_t* __restrict__ out,
                                                           uint32_t slots, uint32_t cap,
                                                           unsigned long long* stats)
{
    for (uint32_t q = blockIdx.x * 256 + threadIdx.x; q < slots; q += gridDim.x * 256)
    {
        const uint64_t b = out[q], e = out[q + 1];
        if (e <= b) continue;
        const unsigned long long c = e - b;
        if (c > cap) atomicAdd(&stats[0], c);
        atomicMax(&stats[1], c);
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

std::size_t lds_cap(const seg::table& t)
{
    std::size_t cap = 1;
    for (int b = 1; b < seg::NBUCKET; ++b)
        if (t.cls[b] != seg::LARGE) cap = std::size_t(1) << b;
    return cap;
}

template <int KIND, bool DESC, bool VALUES>
void run(context& ctx, uint32_t* keys, uint32_t* vals, std::size_t n, const uint64_t* off,
         uint32_t nseg, int key_type, batched_stats* out)
{
    hipStream_t s = ctx.stream();
    seg::table table{};
    const std::vector<int> classes = detail::segmented_sort_table(ctx, key_type, VALUES);
    for (int b = 0; b < seg::NBUCKET; ++b) table.cls[b] = classes[b];
    const uint32_t cap = static_cast<uint32_t>(std::min<std::size_t>(lds_cap(table), 1u << 30));

    // Upper bounds of the device-computed totals.
    const std::size_t samples_max = n / SAMPLE_STRIDE_MIN + 1;
    const std::size_t slots_max = std::size_t(nseg) + n / BUCKET_TARGET + 1;
    const std::size_t tiles_max = n / BTILE + nseg + 1;
    if (slots_max >= (std::size_t(1) << 32) || tiles_max >= (std::size_t(1) << 31))
        throw std::invalid_argument("clc_learned_segmented_sort: input too large");

    layout L;
    const std::size_t o_totals = L.take(64);
    const std::size_t o_stats = L.take(64);
    const std::size_t o_ctr = L.take(64);
    const std::size_t o_hist = L.take(std::size_t(nseg) * 2 * RADIX * 4);
    const std::size_t o_st0 = L.take(tiles_max * RADIX * 4);
    const std::size_t zero_bytes = L.bytes;
    const std::size_t o_st1 = L.take(tiles_max * RADIX * 4);
    const std::size_t o_base = L.take(std::size_t(nseg) * 2 * RADIX * 4);
    const std::size_t o_info = L.take(std::size_t(nseg) * sizeof(seg_info));
    const std::size_t o_soff = L.take((std::size_t(nseg) + 1) * 8);
    const std::size_t o_samples = L.take(samples_max * 4);
    const std::size_t o_models = L.take(std::size_t(nseg) * sizeof(rmi_model));
    const std::size_t o_out = L.take((slots_max + 1) * 8);
    const std::size_t o_tseg = L.take(tiles_max * 4);
    const std::size_t o_kalt = L.take(n * 4);
    const std::size_t o_valt = VALUES ? L.take(n * 4) : 0;
    temp_buffer tmp(ctx.resource(), L.bytes, ctx.stream());
    char* t = tmp.as<char>();
    uint32_t* totals = reinterpret_cast<uint32_t*>(t + o_totals);
    auto* stats = reinterpret_cast<unsigned long long*>(t + o_stats);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(t + o_ctr);
    uint32_t* hist = reinterpret_cast<uint32_t*>(t + o_hist);
    uint32_t* st[2] = { reinterpret_cast<uint32_t*>(t + o_st0),
                        reinterpret_cast<uint32_t*>(t + o_st1) };
    uint32_t* base = reinterpret_cast<uint32_t*>(t + o_base);
    seg_info* info = reinterpret_cast<seg_info*>(t + o_info);
    uint64_t* soff = reinterpret_cast<uint64_t*>(t + o_soff);
    uint32_t* samples = reinterpret_cast<uint32_t*>(t + o_samples);
    rmi_model* models = reinterpret_cast<rmi_model*>(t + o_models);
    uint64_t* bout = reinterpret_cast<uint64_t*>(t + o_out);
    uint32_t* tseg = reinterpret_cast<uint32_t*>(t + o_tseg);
    uint32_t* kalt = reinterpret_cast<uint32_t*>(t + o_kalt);
    uint32_t* valt = VALUES ? reinterpret_cast<uint32_t*>(t + o_valt) : nullptr;

    detail::zero_fill(s, t, zero_bytes);
    {
        timed_launch timer(ctx, "learned_batched_fit");
        hipLaunchKernelGGL(prepare_kernel, dim3(1), dim3(1024), 0, s, off, nseg,
                           static_cast<uint64_t>(n), cap, info, soff, totals);
        hipLaunchKernelGGL(tile_map_kernel, dim3(std::min<uint32_t>(nseg, 4096)), dim3(256), 0,
                           s, info, nseg, tseg);
        hipLaunchKernelGGL((sample_kernel<KIND, DESC>), dim3(nseg), dim3(256), 0, s, keys, info,
                           soff, samples);
        detail::segmented_sort(ctx, samples, TY_U32, nullptr, samples_max, soff, nseg, false);
        hipLaunchKernelGGL(fit_kernel, dim3(nseg), dim3(256), 0, s, samples, soff, info, models,
                           cap, FIT_ERROR_LIMIT, stats + 2);
    }
    const uint32_t tgrid = static_cast<uint32_t>(tiles_max);
    {
        timed_launch timer(ctx, "learned_batched_id_histogram");
        hipLaunchKernelGGL((id_hist_kernel<KIND, DESC>), dim3(tgrid / HU + 1), dim3(1024), 0, s,
                           keys, info, tseg, models, totals, hist);
    }
    hipLaunchKernelGGL(id_scan_kernel, dim3(nseg * 2), dim3(256), 0, s, hist, base);
    const std::size_t lds = std::size_t(BBLOCK / WAVE) * 1024 + RADIX * 4
        + ((sizeof(rmi_model) + 15) & ~std::size_t(15)) + std::size_t(BTILE) * 4
        + std::size_t(BTILE) + 64;
    auto kernel = id_onesweep_kernel<KIND, DESC, VALUES>;
    static bool opted = false;
    if (!opted && lds > 65536)
    {
        CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(lds)));
        opted = true;
    }
    {
        timed_launch timer(ctx, "learned_batched_id_onesweep");
        hipLaunchKernelGGL(kernel, dim3(tgrid), dim3(BBLOCK), lds, s, keys, kalt, vals, valt,
                           info, tseg, models, totals, 0u, base, st[0], st[1], ctr);
    }
    {
        timed_launch timer(ctx, "learned_batched_id_onesweep");
        hipLaunchKernelGGL(kernel, dim3(tgrid), dim3(BBLOCK), lds, s, kalt, keys, valt, vals,
                           info, tseg, models, totals, 1u, base, st[1], nullptr, ctr + 1);
    }
    hipLaunchKernelGGL((copy_back_kernel<VALUES>), dim3(tgrid), dim3(256), 0, s, keys, kalt, vals,
                       valt, info, tseg, totals);
    {
        timed_launch timer(ctx, "learned_batched_bounds");
        hipLaunchKernelGGL((bounds_kernel<KIND, DESC>), dim3(slots_max / 256 + 1), dim3(256), 0,
                           s, keys, info, nseg, models, off, totals,
                           static_cast<uint32_t>(slots_max), bout);
    }
    detail::segmented_sort(ctx, keys, key_type, vals, n, bout, slots_max, DESC);
    if (out != nullptr)
    {
        const uint32_t sg = static_cast<uint32_t>(
            std::max<std::size_t>(1, std::min<std::size_t>(slots_max / 256 + 1, 4096)));
        hipLaunchKernelGGL(bucket_stats_kernel, dim3(sg), dim3(256), 0, s, bout,
                           static_cast<uint32_t>(slots_max), cap, stats);
        unsigned long long h[3];
        uint32_t tot[3];
        CLC_HIP(hipMemcpyAsync(ctx.pinned(), stats, sizeof h, hipMemcpyDeviceToHost, s));
        CLC_HIP(hipStreamSynchronize(s));
        std::memcpy(h, ctx.pinned(), sizeof h);
        CLC_HIP(hipMemcpyAsync(ctx.pinned(), totals, sizeof tot, hipMemcpyDeviceToHost, s));
        CLC_HIP(hipStreamSynchronize(s));
        std::memcpy(tot, ctx.pinned(), sizeof tot);
        // learned segments = those with tiles that passed the check
        out->buckets = tot[1];
        out->sample_size = tot[0];
        out->fit_rejected = static_cast<std::size_t>(h[2]);
        out->fallback_fraction = n ? double(h[0]) / double(n) : 0.0;
        out->max_bucket = static_cast<std::size_t>(h[1]);
    }
    CLC_HIP(hipGetLastError());
}

template <int KIND>
void run_kind(context& ctx, void* keys, void* values, std::size_t n, const uint64_t* off,
              uint32_t nseg, bool desc, int key_type, batched_stats* out)
{
    uint32_t* k = static_cast<uint32_t*>(keys);
    uint32_t* v = static_cast<uint32_t*>(values);
    if (desc)
    {
        if (v) run<KIND, true, true>(ctx, k, v, n, off, nseg, key_type, out);
        else run<KIND, true, false>(ctx, k, v, n, off, nseg, key_type, out);
    }
    else
    {
        if (v) run<KIND, false, true>(ctx, k, v, n, off, nseg, key_type, out);
        else run<KIND, false, false>(ctx, k, v, n, off, nseg, key_type, out);
    }
}

} // namespace
} // namespace learned

namespace detail {

void learned_segmented_sort(context& ctx, void* keys, int key_type, void* values,
                            std::size_t num_items, const uint64_t* offsets,
                            std::size_t num_segments, bool descending,
                            learned::batched_stats* out)
{
    if (out != nullptr) *out = learned::batched_stats{};
    switch (key_type)
    {
        case TY_U32: case TY_I32: case TY_F32: break;
        default:
            throw std::invalid_argument(
                "clc_learned_segmented_sort: keys must be U32, I32 or F32");
    }
    if (num_segments == 0 || num_items == 0) return;
    if (keys == nullptr || offsets == nullptr)
        throw std::invalid_argument("clc_learned_segmented_sort: null keys or offsets");
    if (num_segments >= (std::size_t(1) << 31))
        throw std::invalid_argument("clc_learned_segmented_sort: too many segments");
    ctx.activate();
    const uint32_t nseg = static_cast<uint32_t>(num_segments);
    using namespace learned;
    switch (key_type)
    {
        case TY_U32: run_kind<KEY_UNSIGNED>(ctx, keys, values, num_items, offsets, nseg, descending, key_type, out); break;
        case TY_I32: run_kind<KEY_SIGNED>(ctx, keys, values, num_items, offsets, nseg, descending, key_type, out); break;
        case TY_F32: run_kind<KEY_FLOAT>(ctx, keys, values, num_items, offsets, nseg, descending, key_type, out); break;
    }
}

} // namespace detail
} // namespace clc
