// Learned-bucket sort for gfx950 (no counterpart in the reference, whose
// only sort is detail::radix_sort, Algorithm.cpp:453-527; the plain radix
// sort of this library stays the fallback and the bit-exact reference).
//
//   sample       a strided sample of the keys' order images (as u64; every
//                100th key up to 65,536, every 400th beyond: sample_count),
//                sorted by the plain radix sort
//   fit          a two-level RMI: one root linear model mapping the key to
//                one of 256 leaves (the secant of the sample's CDF between
//                two quantiles), and per leaf a linear model of the
//                empirical CDF (least squares by reduction kernels). Two
//                candidates are fitted at once, the root linear in the key
//                or in its log image (keys spread over many octaves), and
//                the one whose sample predicts less overfill is kept.
//                Slopes are clamped >= 0 and each
//                leaf's bucket range ends where the next leaf's begins, so
//                key -> bucket is a monotone non-decreasing pure function
//                (conversion to double and IEEE rounding are monotone too):
//                equal keys share a bucket and bucket order is key order.
//   check        the sample's predicted buckets estimate how many keys land
//                in buckets too large for LDS that hold more than one
//                distinct key; above a threshold the whole input goes to
//                the plain radix sort (one host read of both candidates' figures)
//   histogram    one read of the keys: counts of the low and high bytes of
//                the 16-bit bucket id (inference fused in)
//   2 passes     stable onesweep passes keyed by the id's low byte, then its
//                high byte (inference fused in): the keys end grouped by
//                bucket, in input order inside a bucket. A direct scatter
//                into 65536 buckets would write one or two keys per bucket
//                per tile; two 256-way passes write long contiguous runs.
//   boundaries   exact bucket offsets by binary search, one thread per bucket
//   finish       the segmented sort over the buckets: buckets that fit LDS
//                are sorted wholly in LDS; larger ones take its onesweep
//                path (constant digit places skipped per bucket, so buckets
//                of one repeated key cost nothing)
//
// The single-scatter form (clc_learned_sort_single_scatter) replaces the
// histogram, the two passes and the boundaries with: per-workgroup bucket
// counts of contiguous chunks (at most 4096 buckets), one exclusive scan of
// the counts bucket-major (exact bucket sizes and each chunk's place in
// every bucket), a copy of the input, and one stable scatter of each
// chunk, tile by tile, sorted by bucket in LDS first. Measured slower than
// the two passes on MI355X (bench learned leg): the scatter writes runs of
// one or two keys per bucket per tile, and at 2^28 keys 4096 buckets hold
// 64Ki keys each, beyond the LDS finish, so the fit check sends the input
// to the plain sort.
#include "context.hpp"
#include "device_common.hpp"
#include "learned_model.hpp"
#include "learned_sort.hpp"
#include "scan_kernels.hpp"
#include "ops.hpp"
#include "radix_tile.hpp"
#include "segmented_sort.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace clc {
namespace learned {
namespace {

using namespace dev;
using namespace radix;

// Onesweep tile of the id passes: 1024 threads x 14 keys with split staging
// and the digits staged as bytes (the largest tile whose LDS, with the
// model and the kernel's static LDS, fits 160 KiB for 64-bit keys).
constexpr int LBLOCK = 1024, LITEMS = 14;
constexpr uint32_t LTILE = LBLOCK * LITEMS;

// ---- sample and fit -------------------------------------------------------------

template <typename K, int KIND, bool DESC>
__global__ __launch_bounds__(256) void sample_kernel(const K* __restrict__ keys, uint32_t m,
                                                     uint32_t every,
                                                     uint64_t* __restrict__ out)
{
    // Four strided reads per thread in flight (each its own cache line).
    constexpr uint32_t U = 4;
    const uint32_t stride = gridDim.x * 256;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < m; i += U * stride)
    {
        K k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            k[u] = keys[size_t(min(i + u * stride, m - 1)) * every + every / 2];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i + u * stride < m)
                out[i + u * stride] = static_cast<uint64_t>(order_bits<K, KIND, DESC>(k[u]));
    }
}

//! Least-squares fits use the moments of (x, y) = (key - xref, index -
//! yref) over units of FIT_UNIT consecutive samples of the sorted sample:
//! one workgroup per unit, then the units of a model combined in a fixed
//! order (deterministic, and no workgroup walks a whole skewed leaf).
//!
//! Two candidate models are fitted side by side (blockIdx.y = c): c = 0 with
//! the root linear in the key, c = 1 with the root linear in its log image.
//! Their leaves are both linear in the key. The host keeps the one whose
//! sample predicts fewer keys in overfull buckets. Per-candidate arrays sit
//! at a fixed stride from each other.
constexpr uint32_t FIT_UNIT = 4096;
constexpr int NMOM = 5; // count, sum x, sum y, sum x^2, sum x*y
constexpr uint32_t CANDIDATES = 2;
//! The root's secant runs through the sample's quantiles 1/ROOT_QUANTILE
//! and 1 - 1/ROOT_QUANTILE.
constexpr uint32_t ROOT_QUANTILE = 512;

//! Moments of samples [lo, hi) of the 256-thread block, x = key - xref
//! and y = index - yref, reduced in a fixed order into out[] (thread 0).
__device__ void unit_moments(const uint64_t* __restrict__ s, uint32_t lo, uint32_t hi,
                             uint64_t xref, uint32_t yref, double* out)
{
    __shared__ double red[4][NMOM];
    double m[NMOM] = { 0, 0, 0, 0, 0 };
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256)
    {
        const double x = static_cast<double>(s[i] - xref); // s[i] >= xref: sorted
        const double y = static_cast<double>(i - yref);
        m[0] += 1.0;
        m[1] += x;
        m[2] += y;
        m[3] += x * x;
        m[4] += x * y;
    }
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < NMOM; ++k)
    {
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) m[k] += __shfl_down(m[k], o, WAVE);
        if (lane == 0) red[wave][k] = m[k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NMOM; ++k) out[k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

//! Sums units [u0, u1) of partial moments in a fixed order (thread 0 gets it).
__device__ void combine_moments(const double* __restrict__ part, uint32_t u0, uint32_t u1,
                                double* out)
{
    __shared__ double red[256][NMOM];
    double m[NMOM] = { 0, 0, 0, 0, 0 };
    for (uint32_t u = u0 + threadIdx.x; u < u1; u += 256)
        for (int k = 0; k < NMOM; ++k) m[k] += part[size_t(u) * NMOM + k];
    for (int k = 0; k < NMOM; ++k) red[threadIdx.x][k] = m[k];
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
            for (int k = 0; k < NMOM; ++k) red[threadIdx.x][k] += red[threadIdx.x + w][k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < NMOM; ++k) out[k] = red[0][k];
}

//! Root model: leaf = a0 * f + b0 maps the key's feature f to 256 leaves
//! by the secant of the sample's empirical CDF through its quantiles q and
//! 1 - q (q = 1/512): samples between them spread over the leaves in
//! proportion to their rank; the tails beyond land in the first and last
//! leaves. A least-squares line over the whole sample instead (fitted to
//! an S-shaped CDF, such as a lognormal's in the log image) clamped about
//! 4% of the keys into each end leaf, where one linear leaf cannot follow
//! the tail, and 3% of the keys of 2^28 lognormal(20, 2) keys went to
//! overfull buckets. All keys at the two quantiles equal: one leaf.
__global__ __launch_bounds__(64) void root_fit_kernel(const uint64_t* __restrict__ s,
                                                      uint32_t m, uint32_t nb,
                                                      rmi_model* models)
{
    const uint32_t c = blockIdx.y;
    if (threadIdx.x != 0) return;
    const uint32_t i1 = m / ROOT_QUANTILE, i2 = m - 1 - i1;
    const uint64_t f1 = c == 1 ? log_image(s[i1]) : s[i1];
    const uint64_t f2 = c == 1 ? log_image(s[i2]) : s[i2];
    const double scale = double(LEAVES) / double(m);
    rmi_model& M = models[c];
    if (f2 > f1)
    {
        M.a0 = double(i2 - i1) * scale / double(f2 - f1);
        M.b0 = double(i1) * scale - M.a0 * static_cast<double>(f1);
    }
    else
    {
        M.a0 = 0.0;
        M.b0 = double(LEAVES / 2);
    }
    M.nb = double(nb);
    M.root_log = c;
}

//! lstart[l] = first sample index whose leaf is >= l (leaves are
//! non-decreasing along the sorted sample); lstart[LEAVES] = m.
__global__ __launch_bounds__(256) void leaf_bounds_kernel(const uint64_t* __restrict__ s,
                                                          uint32_t m, const rmi_model* models,
                                                          uint32_t* __restrict__ lstart_all)
{
    const rmi_model& M = models[blockIdx.y];
    uint32_t* lstart = lstart_all + blockIdx.y * (LEAVES + 1);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256)
    {
        const uint32_t li = leaf_of(M, root_feature(M, s[i]));
        const int lp = i == 0 ? -1 : int(leaf_of(M, root_feature(M, s[i - 1])));
        for (int l = lp + 1; l <= int(li); ++l) lstart[l] = i;
        if (i == m - 1)
            for (uint32_t l = li + 1; l <= LEAVES; ++l) lstart[l] = m;
    }
}

//! Units of each leaf: ustart[l] = first unit of leaf l, ustart[LEAVES] =
//! all units (a leaf of c samples has ceil(c / FIT_UNIT) units).
__global__ __launch_bounds__(256) void leaf_units_kernel(const uint32_t* __restrict__ lstart_all,
                                                         uint32_t* __restrict__ ustart_all)
{
    __shared__ uint32_t tmp[4];
    const uint32_t* lstart = lstart_all + blockIdx.y * (LEAVES + 1);
    uint32_t* ustart = ustart_all + blockIdx.y * (LEAVES + 1);
    const uint32_t l = threadIdx.x;
    const uint32_t cnt = lstart[l + 1] - lstart[l];
    const uint32_t u = (cnt + FIT_UNIT - 1) / FIT_UNIT;
    const uint32_t before = block256_exclusive_add<256>(u, tmp);
    ustart[l] = before;
    if (l == LEAVES - 1) ustart[LEAVES] = before + u;
}

//! Leaf moments: one unit of one leaf, relative to the leaf's first key and
//! index. Launched with an upper bound of units (units_max per candidate);
//! extra blocks return.
__global__ __launch_bounds__(256) void leaf_partial_kernel(const uint64_t* __restrict__ s,
                                                           const uint32_t* __restrict__ lstart_all,
                                                           const uint32_t* __restrict__ ustart_all,
                                                           uint32_t units_max,
                                                           double* __restrict__ part)
{
    const uint32_t* lstart = lstart_all + blockIdx.y * (LEAVES + 1);
    const uint32_t* ustart = ustart_all + blockIdx.y * (LEAVES + 1);
    const uint32_t b = blockIdx.x;
    if (b >= ustart[LEAVES]) return;
    uint32_t lo_l = 0, hi_l = LEAVES; // leaf: last l with ustart[l] <= b
    while (hi_l - lo_l > 1)
    {
        const uint32_t mid = (lo_l + hi_l) / 2;
        if (ustart[mid] <= b) lo_l = mid;
        else hi_l = mid;
    }
    const uint32_t l = lo_l;
    const uint32_t first = lstart[l];
    const uint32_t lo = first + (b - ustart[l]) * FIT_UNIT;
    double mo[NMOM];
    unit_moments(s, lo, min(lstart[l + 1], lo + FIT_UNIT), s[first], first, mo);
    if (threadIdx.x == 0)
        for (int k = 0; k < NMOM; ++k)
            part[(size_t(blockIdx.y) * units_max + b) * NMOM + k] = mo[k];
}

//! Leaf l: linear model of the empirical CDF (slope clamped >= 0) and its
//! bucket range [lo, hi] (proportional to its share of the sample; each
//! range ends where the next begins).
__global__ __launch_bounds__(256) void leaf_fit_kernel(const uint64_t* __restrict__ s,
                                                       uint32_t m, uint32_t nb,
                                                       const uint32_t* __restrict__ lstart_all,
                                                       const uint32_t* __restrict__ ustart_all,
                                                       uint32_t units_max,
                                                       const double* __restrict__ part,
                                                       rmi_model* models)
{
    const uint32_t c = blockIdx.y;
    const uint32_t* lstart = lstart_all + c * (LEAVES + 1);
    const uint32_t* ustart = ustart_all + c * (LEAVES + 1);
    const uint32_t l = blockIdx.x;
    const uint32_t lo = lstart[l], hi = lstart[l + 1];
    double mo[NMOM];
    combine_moments(part + size_t(c) * units_max * NMOM, ustart[l], ustart[l + 1], mo);
    if (threadIdx.x == 0)
    {
        double a = 0, b = (double(lo) + double(hi)) / (2.0 * m);
        if (hi - lo >= 2)
        {
            double ai, ci;
            line_from_moments(mo, ai, ci); // index - lo ~ ai * (x - s[lo]) + ci
            a = ai / double(m);
            b = (double(lo) + ci - ai * static_cast<double>(s[lo])) / double(m);
        }
        const uint32_t blo = static_cast<uint32_t>(uint64_t(nb) * lo / m);
        const uint32_t next =
            l + 1 < LEAVES ? static_cast<uint32_t>(uint64_t(nb) * hi / m) : nb - 1;
        rmi_leaf& L = models[c].leaf[l];
        L.a = a * double(nb); // positions in buckets, not CDF fractions
        L.b = b * double(nb);
        L.lo = min(blo, nb - 1);
        L.hi = max(L.lo, min(next, nb - 1));
    }
}

//! Per predicted bucket of the sample: its first index (buckets are
//! non-decreasing along the sorted sample); nb entries per candidate.
__global__ __launch_bounds__(256) void sample_runs_kernel(const uint64_t* __restrict__ s,
                                                          uint32_t m, uint32_t nb,
                                                          const rmi_model* models,
                                                          uint32_t* __restrict__ runs_all)
{
    __shared__ rmi_model sm;
    load_model(&sm, models + blockIdx.y);
    __syncthreads();
    uint32_t* run_start = runs_all + size_t(blockIdx.y) * nb;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256)
    {
        const uint32_t b = bucket_of(sm, s[i]);
        if (i == 0 || bucket_of(sm, s[i - 1]) != b) run_start[b] = i;
    }
}

//! Per candidate c: quality[2c] = largest predicted bucket (sample keys);
//! quality[2c + 1] = sample keys in buckets whose estimated size exceeds cap
//! and that hold more than one distinct key (those would go through the
//! onesweep fallback). Reduced per block, one atomic per block and value.
__global__ __launch_bounds__(256) void sample_quality_kernel(
    const uint64_t* __restrict__ s, uint32_t m, uint32_t nb, const rmi_model* models,
    const uint32_t* __restrict__ runs_all, double keys_per_sample, uint32_t cap,
    unsigned long long* quality)
{
    __shared__ rmi_model sm;
    __shared__ unsigned long long s_max[4], s_sum[4];
    load_model(&sm, models + blockIdx.y);
    __syncthreads();
    const uint32_t* run_start = runs_all + size_t(blockIdx.y) * nb;
    unsigned long long mx = 0, sum = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256)
    {
        const uint32_t b = bucket_of(sm, s[i]);
        if (i + 1 < m && bucket_of(sm, s[i + 1]) == b) continue; // not a run end
        const uint32_t first = run_start[b];
        const uint32_t cnt = i + 1 - first;
        mx = max(mx, static_cast<unsigned long long>(cnt));
        if (s[first] != s[i] && double(cnt) * keys_per_sample > double(cap)) sum += cnt;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
    {
        mx = max(mx, static_cast<unsigned long long>(__shfl_xor(mx, o, WAVE)));
        sum += __shfl_xor(sum, o, WAVE);
    }
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    if (lane == 0)
    {
        s_max[wave] = mx;
        s_sum[wave] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        mx = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
        sum = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
        if (mx != 0) atomicMax(&quality[2 * blockIdx.y], mx);
        if (sum != 0) atomicAdd(&quality[2 * blockIdx.y + 1], sum);
    }
}

// ---- the fused-inference passes ---------------------------------------------------

template <typename K, int KIND, bool DESC>
__global__ __launch_bounds__(256) void id_histogram_kernel(const K* __restrict__ keys,
                                                           uint32_t n, const rmi_model* model,
                                                           uint32_t* __restrict__ hist)
{
    // C interleaved copies per bin (copy = lane mod C): skewed keys put many
    // lanes on one bucket, which would serialise 64 deep on a single copy.
    constexpr uint32_t C = 8;
    __shared__ rmi_model sm;
    __shared__ uint32_t s_h[2 * RADIX * C];
    load_model(&sm, model);
    for (uint32_t i = threadIdx.x; i < 2 * RADIX * C; i += 256) s_h[i] = 0;
    __syncthreads();
    const uint32_t copy = threadIdx.x % C;
    auto count = [&](K key) {
        const uint32_t b = bucket_of(sm, static_cast<uint64_t>(order_bits<K, KIND, DESC>(key)));
        atomicAdd(&s_h[(b & 0xFFu) * C + copy], 1u);
        atomicAdd(&s_h[(RADIX + (b >> 8)) * C + copy], 1u);
    };
    const uint32_t stride = gridDim.x * 256;
    uint32_t start = 0;
    if (reinterpret_cast<uintptr_t>(keys) % 16 == 0)
    {
        // Two 16-byte loads per thread in flight (streamed: read once here),
        // clamped to the last vector; then the keys they hold.
        constexpr uint32_t PER = 16 / sizeof(K);
        const uint32_t nvec = n / PER;
        const uint4* v = reinterpret_cast<const uint4*>(keys);
        for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nvec; i += 2 * stride)
        {
            const uint4 q0 = load_stream(v + i);
            const uint4 q1 = load_stream(v + min(i + stride, nvec - 1));
            const K* k0 = reinterpret_cast<const K*>(&q0);
            const K* k1 = reinterpret_cast<const K*>(&q1);
#pragma unroll
            for (uint32_t j = 0; j < PER; ++j) count(k0[j]);
            if (i + stride < nvec)
#pragma unroll
                for (uint32_t j = 0; j < PER; ++j) count(k1[j]);
        }
        start = nvec * PER;
    }
    // The rest (or an unaligned array): U independent loads per thread in
    // flight before any inference.
    constexpr uint32_t U = 4;
    for (uint32_t i = start + blockIdx.x * 256 + threadIdx.x; i < n; i += U * stride)
    {
        K k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) k[u] = keys[min(i + u * stride, n - 1)]; // unconditional
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
        {
            if (i + u * stride >= n) break;
            count(k[u]);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 2 * RADIX; i += 256)
    {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t j = 0; j < C; ++j) c += s_h[i * C + j];
        if (c != 0) atomicAdd(&hist[i], c);
    }
}

__global__ __launch_bounds__(256) void id_scan_kernel(const uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ base)
{
    __shared__ uint32_t tmp[4];
    const uint32_t i = blockIdx.x * RADIX + threadIdx.x;
    base[i] = block256_exclusive_add<256>(hist[i], tmp);
}

template <typename K, int KIND, bool DESC, bool VALUES>
__global__ __launch_bounds__(LBLOCK) void id_onesweep_kernel(
    const K* __restrict__ keys_in, K* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out, uint32_t n,
    uint32_t shift, const rmi_model* model, const uint32_t* __restrict__ digit_base,
    uint32_t* status, uint32_t* status_next, uint32_t* tile_counter)
{
    constexpr int W = LBLOCK / WAVE;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem);
    uint32_t* s_gbase = s_hist + W * RADIX;
    rmi_model* s_model = reinterpret_cast<rmi_model*>(s_gbase + RADIX);
    K* s_keys = reinterpret_cast<K*>(reinterpret_cast<unsigned char*>(s_model)
                                     + ((sizeof(rmi_model) + 15) & ~size_t(15)));
    uint8_t* s_dig = reinterpret_cast<uint8_t*>(s_keys + LTILE);
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_tile = fetch_add_agent(tile_counter, 1u);
    for (uint32_t i = tid; i < W * RADIX; i += LBLOCK) s_hist[i] = 0;
    load_model(s_model, model);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (status_next != nullptr && tid < RADIX)
        status_next[static_cast<size_t>(tile) * RADIX + tid] = 0;
    const id_digit<K, KIND, DESC> dig{ s_model, shift, s_dig };
    // Split staging: values pass through the key buffer after the keys.
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
    if (tile * LTILE + LTILE <= n)
        onesweep_tile<K, uint32_t, VALUES, LBLOCK, LITEMS, true, true>(
            keys_in, keys_out, vals_in, vals_out, n, dig, digit_base, status, tile, s_hist,
            s_keys, s_vals, s_gbase, s_tmp);
    else
        onesweep_tile<K, uint32_t, VALUES, LBLOCK, LITEMS, false, true>(
            keys_in, keys_out, vals_in, vals_out, n, dig, digit_base, status, tile, s_hist,
            s_keys, s_vals, s_gbase, s_tmp);
}

//! Exact bucket offsets of keys already grouped by non-decreasing bucket:
//! off[c] = first index whose bucket is >= c, by binary search (one thread
//! per bucket, about log2(n) dependent reads each), off[nb] = n.
template <typename K, int KIND, bool DESC>
__global__ __launch_bounds__(256) void boundary_kernel(const K* __restrict__ keys, uint32_t n,
                                                       const rmi_model* model, uint32_t nb,
                                                       uint64_t* __restrict__ off)
{
    __shared__ rmi_model sm;
    load_model(&sm, model);
    __syncthreads();
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c > nb) return;
    uint32_t lo = 0, hi = n; // answer in [lo, hi]
    if (c == nb) lo = n;
    while (lo < hi)
    {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (bucket_of(sm, static_cast<uint64_t>(order_bits<K, KIND, DESC>(keys[mid]))) >= c)
            hi = mid;
        else
            lo = mid + 1;
    }
    off[c] = lo;
}

//! stats[0] = largest bucket, stats[1] = keys in buckets above cap.
__global__ __launch_bounds__(256) void bucket_stats_kernel(const uint64_t* __restrict__ off,
                                                           uint32_t nb, uint32_t cap,
                                                           unsigned long long* stats)
{
    for (uint32_t b = blockIdx.x * 256 + threadIdx.x; b < nb; b += gridDim.x * 256)
    {
        const unsigned long long c = off[b + 1] - off[b];
        atomicMax(&stats[0], c);
        if (c > cap) atomicAdd(&stats[1], c);
    }
}

// ---- the single fused scatter (at most 4096 buckets) ------------------------------

//! Tile of the single scatter's workgroups, and its largest bucket count.
constexpr int SBLOCK = 1024, SITEMS = 6;
constexpr uint32_t STILE = SBLOCK * SITEMS;
constexpr uint32_t SS_MAX_BUCKETS = 4096;

//! Dynamic LDS of ss_scatter_kernel: per-wave digit counters, the running
//! bucket offsets, each bucket's first sorted slot in the tile, the packed
//! words, the tile's values and keys, and the model.
constexpr std::size_t ss_lds(std::size_t kbytes, bool values)
{
    return std::size_t(SBLOCK / WAVE) * RADIX * 4 + 2 * SS_MAX_BUCKETS * 4 + STILE * 4
        + (values ? STILE * 4 : 0) + STILE * kbytes + sizeof(rmi_model);
}

//! Workgroup w of the grid counts the buckets of its chunk [w * chunk,
//! min(n, (w + 1) * chunk)) with the inference fused in, and stores them
//! bucket-major, counts[b * grid + w]: one exclusive scan over the array
//! then gives every (bucket, workgroup) pair its first output slot, so the
//! bucket sizes and every workgroup's place inside each bucket are exact.
template <typename K, int KIND, bool DESC>
__global__ __launch_bounds__(1024) void ss_count_kernel(const K* __restrict__ keys, uint32_t n,
                                                        uint32_t chunk, const rmi_model* model,
                                                        uint32_t nb, uint32_t* __restrict__ counts)
{
    __shared__ rmi_model sm;
    __shared__ uint32_t s_h[SS_MAX_BUCKETS];
    const uint32_t tid = threadIdx.x;
    load_model(&sm, model);
    for (uint32_t b = tid; b < nb; b += 1024) s_h[b] = 0;
    __syncthreads();
    const uint32_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    constexpr uint32_t U = 4; // loads in flight per thread
    for (uint32_t i0 = lo; i0 < hi; i0 += U * 1024)
    {
        K k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) k[u] = keys[min(i0 + u * 1024 + tid, hi - 1)];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (i0 + u * 1024 + tid < hi)
            {
                const uint32_t b = bucket_of(sm, static_cast<uint64_t>(order_bits<K, KIND, DESC>(k[u])));
                atomicAdd(&s_h[min(b, nb - 1)], 1u);
            }
    }
    __syncthreads();
    for (uint32_t b = tid; b < nb; b += 1024)
        counts[size_t(b) * gridDim.x + blockIdx.x] = s_h[b];
}

//! The single fused scatter: workgroup w walks its chunk in tiles of STILE
//! keys, in order. Per tile: each key's bucket (inference), a stable sort
//! of the words (bucket << 16 | position) in LDS by the bucket's id_bits
//! bits (lds_passes), then each key goes to its bucket's running offset
//! plus its rank among the tile's keys of that bucket, and the offsets
//! advance by the tile's counts. Keys of one bucket keep their input order
//! (tile order inside a workgroup, workgroup order from the scan), so the
//! scatter is stable.
template <typename K, int KIND, bool DESC, bool VALUES>
__global__ __launch_bounds__(SBLOCK) void ss_scatter_kernel(
    const K* __restrict__ keys_in, K* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out, uint32_t n,
    uint32_t chunk, const rmi_model* model, uint32_t nb, uint32_t id_bits,
    const uint32_t* __restrict__ offsets)
{
    constexpr int W = SBLOCK / WAVE;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* s_hist = reinterpret_cast<uint32_t*>(smem);
    uint32_t* s_off = s_hist + W * RADIX;
    uint32_t* s_first = s_off + SS_MAX_BUCKETS;
    uint32_t* s_w = s_first + SS_MAX_BUCKETS;
    uint32_t* s_vals = s_w + STILE;
    K* s_keys = reinterpret_cast<K*>(s_vals + (VALUES ? STILE : 0));
    rmi_model* s_model = reinterpret_cast<rmi_model*>(s_keys + STILE);
    __shared__ uint32_t s_tmp[4];
    const uint32_t tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    for (uint32_t i = tid; i < W * RADIX; i += SBLOCK) s_hist[i] = 0;
    for (uint32_t b = tid; b < nb; b += SBLOCK)
        s_off[b] = offsets[size_t(b) * gridDim.x + blockIdx.x];
    load_model(s_model, model);
    __syncthreads();
    const uint32_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    // Position of row r of this thread (the layout lds_pass expects).
    const uint32_t row0 = wave * (WAVE * SITEMS) + lane;
    for (uint32_t t0 = lo; t0 < hi; t0 += STILE)
    {
        const uint32_t m = min(STILE, hi - t0);
        uint32_t w[SITEMS], unused[SITEMS];
#pragma unroll
        for (int r = 0; r < SITEMS; ++r)
        {
            const uint32_t p = row0 + r * WAVE;
            w[r] = 0;
            if (p < m)
            {
                const K k = keys_in[t0 + p];
                s_keys[p] = k;
                if constexpr (VALUES) s_vals[p] = vals_in[t0 + p];
                const uint32_t b =
                    bucket_of(*s_model, static_cast<uint64_t>(order_bits<K, KIND, DESC>(k)));
                w[r] = (min(b, nb - 1) << 16) | p;
            }
        }
        lds_passes<uint32_t, 0, SBLOCK, SITEMS>(w, unused, 16, id_bits, 8, m, SITEMS, s_hist,
                                                s_w, nullptr, s_tmp);
        // s_w now holds the tile's words in sorted order (slots below m),
        // and w[r] the word at slot row0 + r * 64. First slot of each run.
#pragma unroll
        for (int r = 0; r < SITEMS; ++r)
        {
            const uint32_t i = row0 + r * WAVE;
            if (i < m)
            {
                const uint32_t id = w[r] >> 16;
                if (i == 0 || (s_w[i - 1] >> 16) != id) s_first[id] = i;
            }
        }
        __syncthreads();
        uint32_t run_len[SITEMS];
#pragma unroll
        for (int r = 0; r < SITEMS; ++r)
        {
            const uint32_t i = row0 + r * WAVE;
            run_len[r] = 0;
            if (i < m)
            {
                const uint32_t id = w[r] >> 16, pos = w[r] & 0xFFFFu;
                const uint32_t first = s_first[id];
                const uint32_t dest = s_off[id] + (i - first);
                keys_out[dest] = s_keys[pos];
                if constexpr (VALUES) vals_out[dest] = s_vals[pos];
                if (i + 1 == m || (s_w[i + 1] >> 16) != id) run_len[r] = i + 1 - first;
            }
        }
        __syncthreads(); // every offset read before any advances
#pragma unroll
        for (int r = 0; r < SITEMS; ++r)
            if (run_len[r] != 0) s_off[w[r] >> 16] += run_len[r];
        __syncthreads(); // offsets advanced; the tile's LDS is free again
    }
}

//! Bucket offsets of the single scatter: off[b] = first slot of bucket b
//! (its first workgroup's offset), off[nb] = n.
__global__ __launch_bounds__(256) void ss_bounds_kernel(const uint32_t* __restrict__ offsets,
                                                        uint32_t groups, uint32_t nb, uint32_t n,
                                                        uint64_t* __restrict__ off)
{
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b < nb) off[b] = offsets[size_t(b) * groups];
    else if (b == nb) off[b] = n;
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

//! Largest segment the segmented sort keeps in LDS under table t.
std::size_t lds_capacity(const seg::table& t)
{
    std::size_t cap = 1;
    for (int b = 1; b < seg::NBUCKET; ++b)
        if (t.cls[b] != seg::LARGE) cap = std::size_t(1) << b;
    return cap;
}

template <typename K, int KIND, bool DESC, bool VALUES>
void run(context& ctx, K* keys, uint32_t* vals, uint32_t n, int key_type, bool single,
         stats* out)
{
    hipStream_t s = ctx.stream();
    const uint32_t m = sample_count(n);
    // Buckets: about 4096 keys each, 512 to 65536 (two 8-bit passes), or at
    // most 4096 for the single scatter.
    const uint32_t nb_max = single ? SS_MAX_BUCKETS : 65536u;
    uint32_t nb = 512;
    while (nb < nb_max && nb * 4096u < n) nb <<= 1;
    seg::table table{};
    const std::vector<int> classes = detail::segmented_sort_table(ctx, key_type, VALUES);
    for (int b = 0; b < seg::NBUCKET; ++b) table.cls[b] = classes[b];
    const std::size_t cap = lds_capacity(table);
    const uint32_t tiles = (n + LTILE - 1) / LTILE;

    constexpr uint32_t NC = CANDIDATES;
    layout L;
    const std::size_t o_quality = L.take(NC * 2 * 8);
    const std::size_t o_stats = L.take(64);
    const std::size_t o_hist = L.take(2 * RADIX * 4);
    const std::size_t o_ctr = L.take(64);
    const std::size_t o_runs = L.take(std::size_t(NC) * nb * 4);
    const std::size_t o_st0 = L.take(std::size_t(tiles) * RADIX * 4);
    const std::size_t zero_bytes = L.bytes;
    const std::size_t o_st1 = L.take(std::size_t(tiles) * RADIX * 4);
    const std::size_t o_model = L.take(NC * sizeof(rmi_model));
    const std::size_t o_lstart = L.take(NC * (LEAVES + 1) * 4);
    const std::size_t o_ustart = L.take(NC * (LEAVES + 1) * 4);
    const uint32_t sample_units = (m + FIT_UNIT - 1) / FIT_UNIT;
    const uint32_t leaf_units_max = LEAVES + sample_units; // a leaf adds at most one partial unit
    const std::size_t o_lpart = L.take(std::size_t(NC) * leaf_units_max * NMOM * 8);
    const std::size_t o_base = L.take(2 * RADIX * 4);
    const std::size_t o_sample = L.take(std::size_t(m) * 8);
    const std::size_t o_off = L.take((std::size_t(nb) + 1) * 8);
    const std::size_t o_kalt = L.take(std::size_t(n) * sizeof(K));
    const std::size_t o_valt = VALUES ? L.take(std::size_t(n) * 4) : 0;
    temp_buffer tmp(ctx.resource(), L.bytes, ctx.stream());
    char* t = tmp.as<char>();
    auto* quality = reinterpret_cast<unsigned long long*>(t + o_quality);
    auto* bstats = reinterpret_cast<unsigned long long*>(t + o_stats);
    uint32_t* hist = reinterpret_cast<uint32_t*>(t + o_hist);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(t + o_ctr);
    uint32_t* runs = reinterpret_cast<uint32_t*>(t + o_runs);
    uint32_t* st[2] = { reinterpret_cast<uint32_t*>(t + o_st0),
                        reinterpret_cast<uint32_t*>(t + o_st1) };
    rmi_model* models = reinterpret_cast<rmi_model*>(t + o_model);
    uint32_t* lstart = reinterpret_cast<uint32_t*>(t + o_lstart);
    uint32_t* ustart = reinterpret_cast<uint32_t*>(t + o_ustart);
    double* lpart = reinterpret_cast<double*>(t + o_lpart);
    uint32_t* base = reinterpret_cast<uint32_t*>(t + o_base);
    uint64_t* sample = reinterpret_cast<uint64_t*>(t + o_sample);
    uint64_t* off = reinterpret_cast<uint64_t*>(t + o_off);
    K* kalt = reinterpret_cast<K*>(t + o_kalt);
    uint32_t* valt = VALUES ? reinterpret_cast<uint32_t*>(t + o_valt) : nullptr;
    const uint32_t grid = std::max(1u, std::min<uint32_t>((n + 2047) / 2048,
                                                          uint32_t(ctx.facts().cu_count) * 8));

    detail::zero_fill(s, t, zero_bytes);
    {
        timed_launch timer(ctx, "learned_sample_fit");
        hipLaunchKernelGGL((sample_kernel<K, KIND, DESC>), dim3(grid), dim3(256), 0, s, keys,
                           m, n / m, sample);
        detail::radix_sort(ctx, sample, TY_U64, nullptr, m, false);
        // Both candidate models at once: blockIdx.y selects the candidate.
        const uint32_t sgrid = std::max(1u, std::min<uint32_t>((m + 2047) / 2048,
                                                               uint32_t(ctx.facts().cu_count) * 4));
        hipLaunchKernelGGL(root_fit_kernel, dim3(1, NC), dim3(64), 0, s, sample, m, nb, models);
        hipLaunchKernelGGL(leaf_bounds_kernel, dim3(sgrid, NC), dim3(256), 0, s, sample, m,
                           models, lstart);
        hipLaunchKernelGGL(leaf_units_kernel, dim3(1, NC), dim3(256), 0, s, lstart, ustart);
        hipLaunchKernelGGL(leaf_partial_kernel, dim3(leaf_units_max, NC), dim3(256), 0, s,
                           sample, lstart, ustart, leaf_units_max, lpart);
        hipLaunchKernelGGL(leaf_fit_kernel, dim3(LEAVES, NC), dim3(256), 0, s, sample, m, nb,
                           lstart, ustart, leaf_units_max, lpart, models);
        hipLaunchKernelGGL(sample_runs_kernel, dim3(sgrid, NC), dim3(256), 0, s, sample, m, nb,
                           models, runs);
        hipLaunchKernelGGL(sample_quality_kernel, dim3(sgrid, NC), dim3(256), 0, s, sample, m,
                           nb, models, runs, double(n) / double(m), uint32_t(cap), quality);
    }
    unsigned long long q[2 * NC];
    CLC_HIP(hipMemcpyAsync(ctx.pinned(), quality, sizeof q, hipMemcpyDeviceToHost, s));
    CLC_HIP(hipStreamSynchronize(s));
    std::memcpy(q, ctx.pinned(), sizeof q);
    // The candidate with fewer sample keys in overfull buckets, then the
    // smaller largest bucket; the linear root on a tie.
    uint32_t pick = 0;
    for (uint32_t c = 1; c < NC; ++c)
        if (q[2 * c + 1] < q[2 * pick + 1]
            || (q[2 * c + 1] == q[2 * pick + 1] && q[2 * c] < q[2 * pick]))
            pick = c;
    const rmi_model* model = models + pick;
    const double fit_error = double(q[2 * pick + 1]) / double(m);
    if (out != nullptr)
    {
        out->num_buckets = nb;
        out->sample_size = m;
        out->sample_fit_error = fit_error;
        out->root_log = static_cast<int>(pick);
    }
    if (fit_error > FIT_ERROR_LIMIT)
    {
        // The model would send too much to the fallback: plain sort for all.
        detail::radix_sort(ctx, keys, key_type, vals, n, DESC);
        if (out != nullptr)
        {
            out->used_model = 0;
            out->max_bucket_overfill = 0;
            out->fallback_fraction = 1.0;
        }
        return;
    }

    if (single)
    {
        // Workgroups of one resident tile each (the scatter's LDS holds one
        // per CU), chunks of whole tiles.
        const uint32_t groups_want = std::max(1u, uint32_t(ctx.facts().cu_count));
        const uint32_t chunk =
            std::max(1u, (n + groups_want * STILE - 1) / (groups_want * STILE)) * STILE;
        const uint32_t groups = (n + chunk - 1) / chunk;
        const std::size_t cells = std::size_t(nb) * groups;
        temp_buffer cnt(ctx.resource(), 2 * cells * 4, ctx.stream());
        uint32_t* counts = cnt.as<uint32_t>();
        uint32_t* offsets = counts + cells;
        uint32_t id_bits = 0;
        while ((1u << id_bits) < nb) ++id_bits;
        {
            timed_launch timer(ctx, "learned_single_scatter_count");
            hipLaunchKernelGGL((ss_count_kernel<K, KIND, DESC>), dim3(groups), dim3(1024), 0, s,
                               keys, n, chunk, model, nb, counts);
        }
        const uint32_t zero = 0;
        detail::scan(ctx, counts, offsets, cells, TY_U32, OP_PLUS, &zero, true);
        // The scatter reads the keys from a copy and writes them to their
        // buckets in place of the input.
        CLC_HIP(hipMemcpyAsync(kalt, keys, std::size_t(n) * sizeof(K), hipMemcpyDeviceToDevice, s));
        if (VALUES)
            CLC_HIP(hipMemcpyAsync(valt, vals, std::size_t(n) * 4, hipMemcpyDeviceToDevice, s));
        const std::size_t slds = ss_lds(sizeof(K), VALUES);
        auto sk = ss_scatter_kernel<K, KIND, DESC, VALUES>;
        static bool s_opted = false;
        if (!s_opted)
        {
            CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(sk),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(slds)));
            s_opted = true;
        }
        {
            timed_launch timer(ctx, "learned_single_scatter");
            hipLaunchKernelGGL(sk, dim3(groups), dim3(SBLOCK), slds, s, kalt, keys, valt, vals, n,
                               chunk, model, nb, id_bits, offsets);
        }
        hipLaunchKernelGGL(ss_bounds_kernel, dim3(nb / 256 + 1), dim3(256), 0, s, offsets,
                           groups, nb, n, off);
    }
    else
    {
        {
            timed_launch timer(ctx, "learned_id_histogram");
            hipLaunchKernelGGL((id_histogram_kernel<K, KIND, DESC>), dim3(grid), dim3(256), 0, s,
                               keys, n, model, hist);
        }
        hipLaunchKernelGGL(id_scan_kernel, dim3(2), dim3(256), 0, s, hist, base);
        const std::size_t lds = std::size_t(LBLOCK / WAVE) * 1024 + RADIX * 4
            + ((sizeof(rmi_model) + 15) & ~std::size_t(15)) + std::size_t(LTILE) * sizeof(K)
            + std::size_t(LTILE) + 64;
        auto kernel = id_onesweep_kernel<K, KIND, DESC, VALUES>;
        static bool opted = false;
        if (!opted && lds > 65536)
        {
            CLC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
            opted = true;
        }
        {
            timed_launch timer(ctx, "learned_id_onesweep");
            hipLaunchKernelGGL(kernel, dim3(tiles), dim3(LBLOCK), lds, s, keys, kalt, vals, valt, n,
                               0u, model, base, st[0], st[1], ctr);
        }
        {
            timed_launch timer(ctx, "learned_id_onesweep");
            hipLaunchKernelGGL(kernel, dim3(tiles), dim3(LBLOCK), lds, s, kalt, keys, valt, vals, n,
                               8u, model, base + RADIX, st[1], nullptr, ctr + 1);
        }
        {
            timed_launch timer(ctx, "learned_boundaries");
            hipLaunchKernelGGL((boundary_kernel<K, KIND, DESC>), dim3(nb / 256 + 1), dim3(256), 0,
                               s, keys, n, model, nb, off);
        }
    }
    detail::segmented_sort(ctx, keys, key_type, vals, n, off, nb, DESC);
    if (out != nullptr)
    {
        hipLaunchKernelGGL(bucket_stats_kernel, dim3(std::max(1u, nb / 256)), dim3(256), 0, s,
                           off, nb, uint32_t(cap), bstats);
        CLC_HIP(hipMemcpyAsync(ctx.pinned(), bstats, 2 * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s));
        CLC_HIP(hipStreamSynchronize(s));
        unsigned long long b2[2];
        std::memcpy(b2, ctx.pinned(), sizeof b2);
        out->used_model = 1;
        out->max_bucket_overfill = double(b2[0]) / (double(n) / double(nb));
        out->fallback_fraction = double(b2[1]) / double(n);
    }
    CLC_HIP(hipGetLastError());
}

template <typename K, int KIND>
void run_typed(context& ctx, void* keys, void* values, uint32_t n, bool desc, int key_type,
               bool single, stats* out)
{
    K* k = static_cast<K*>(keys);
    uint32_t* v = static_cast<uint32_t*>(values);
    if (desc)
    {
        if (v) run<K, KIND, true, true>(ctx, k, v, n, key_type, single, out);
        else run<K, KIND, true, false>(ctx, k, v, n, key_type, single, out);
    }
    else
    {
        if (v) run<K, KIND, false, true>(ctx, k, v, n, key_type, single, out);
        else run<K, KIND, false, false>(ctx, k, v, n, key_type, single, out);
    }
}

} // namespace
} // namespace learned

namespace detail {

void learned_sort(context& ctx, void* keys, int key_type, void* values, std::size_t n,
                  bool descending, bool single_scatter, learned::stats* out)
{
    switch (key_type)
    {
        case TY_U32: case TY_I32: case TY_F32: case TY_U64: case TY_I64: case TY_F64: break;
        default:
            throw std::invalid_argument(
                "clc_learned_sort: keys must be I32, U32, F32, I64, U64 or F64");
    }
    if (out != nullptr) *out = learned::stats{};
    if (n <= 1) return;
    check_size(n);
    if (keys == nullptr) throw std::invalid_argument("clc_learned_sort: null keys");
    ctx.activate();
    // Too small for a model to pay off, or too large for 30-bit look-back
    // counts: the plain sort does it.
    if (n < learned::MIN_KEYS || n >= (1u << 30))
    {
        radix_sort(ctx, keys, key_type, values, n, descending);
        if (out != nullptr) out->fallback_fraction = 1.0;
        return;
    }
    const uint32_t m = static_cast<uint32_t>(n);
    using namespace learned;
    switch (key_type)
    {
        case TY_U32: run_typed<uint32_t, KEY_UNSIGNED>(ctx, keys, values, m, descending, key_type, single_scatter, out); break;
        case TY_I32: run_typed<uint32_t, KEY_SIGNED>(ctx, keys, values, m, descending, key_type, single_scatter, out); break;
        case TY_F32: run_typed<uint32_t, KEY_FLOAT>(ctx, keys, values, m, descending, key_type, single_scatter, out); break;
        case TY_U64: run_typed<uint64_t, KEY_UNSIGNED>(ctx, keys, values, m, descending, key_type, single_scatter, out); break;
        case TY_I64: run_typed<uint64_t, KEY_SIGNED>(ctx, keys, values, m, descending, key_type, single_scatter, out); break;
        case TY_F64: run_typed<uint64_t, KEY_FLOAT>(ctx, keys, values, m, descending, key_type, single_scatter, out); break;
    }
}

} // namespace detail
} // namespace clc
