// The learned CDF model shared by the learned-bucket sorts (single array
// and batched segments): a two-level RMI, one root linear model over a
// feature of the key choosing one of 256 leaves, one linear CDF model per
// leaf, and the monotone key -> bucket map it defines.
#pragma once

#include "device_common.hpp"
#include "radix_tile.hpp"

#include <cstdint>

namespace clc {
namespace learned {

constexpr uint32_t LEAVES = 256;
//! The sample of m keys is every (m / count)-th key, from the middle of the
//! first stride, count = sample_count(m): every 100th key up to 65,536
//! samples, and beyond that every 400th once that gives more. A larger
//! sample costs sorting and fitting time without a better model: the 2^28
//! u64 lognormal sort took 6.40 / 6.33 / 6.24 / 6.24 ms with every 100th /
//! 200th / 400th / 800th key (keys in overfull buckets 0.30 / 0.32 / 0.54 /
//! 0.80%).
constexpr uint32_t SAMPLE_STRIDE_MIN = 100;
constexpr uint32_t SAMPLE_STRIDE_MAX = 400;
constexpr uint32_t SAMPLE_FLOOR = 65536;
__host__ __device__ inline uint32_t sample_count(uint32_t m)
{
    const uint32_t dense = m / SAMPLE_STRIDE_MIN, sparse = m / SAMPLE_STRIDE_MAX;
    const uint32_t capped = dense < SAMPLE_FLOOR ? dense : SAMPLE_FLOOR;
    return sparse > capped ? sparse : capped;
}

struct rmi_leaf
{
    double a, b; // bucket position ~ a * x + b (the CDF times the bucket count)
    uint32_t lo, hi; // bucket range of the leaf
};

struct rmi_model
{
    double a0, b0; // leaf ~ a0 * f + b0, f = root_feature(x)
    double nb; // number of buckets
    uint32_t root_log; // root feature: 0 the key, 1 its log image
    uint32_t pad;
    rmi_leaf leaf[LEAVES];
};

//! Log image of an order image: the bits of double(x) read as an integer,
//! a monotone non-decreasing, piecewise linear approximation of
//! 2^52 * (log2(x) + 1023). A root linear in it spreads keys that span
//! many octaves (lognormal, Zipf) over the leaves by magnitude.
__device__ __forceinline__ uint64_t log_image(uint64_t x)
{
    return static_cast<uint64_t>(__double_as_longlong(static_cast<double>(x)));
}

__device__ __forceinline__ double root_feature(const rmi_model& m, uint64_t bits)
{
    return static_cast<double>(m.root_log ? log_image(bits) : bits);
}

//! Leaf of root feature f: a0 * f + b0 clamped to [0, LEAVES - 1] and
//! truncated (clamping first keeps the conversion in range).
__device__ __forceinline__ uint32_t leaf_of(const rmi_model& m, double f)
{
    const double r = fmin(fmax(fma(m.a0, f, m.b0), 0.0), double(LEAVES - 1));
    return static_cast<uint32_t>(r);
}

//! Monotone non-decreasing bucket of an order image x: the root feature,
//! its conversion to double and every model step are monotone, leaf slopes
//! are >= 0 and leaf bucket ranges are ordered. The position a * x + b is
//! truncated (clamped to the u32 range first) and clamped to the leaf's
//! [lo, hi] in integers, which equals clamping it in doubles first: for
//! integers lo <= hi, trunc(p) <= lo when p <= lo and >= hi when p >= hi.
__device__ __forceinline__ uint32_t bucket_of(const rmi_model& m, uint64_t bits)
{
    const double x = static_cast<double>(bits);
    const rmi_leaf& L = m.leaf[leaf_of(m, root_feature(m, bits))];
    const double p = fmin(fmax(fma(L.a, x, L.b), 0.0), 4294967295.0);
    return min(max(static_cast<uint32_t>(p), L.lo), L.hi);
}

//! Onesweep digit functor: byte shift/8 of the key's bucket under model m.
template <typename K, int KIND, bool DESC> struct id_digit
{
    //! Inference is costly: one evaluation per key and pass, the digit
    //! staged in s_dig for the write-out.
    static constexpr bool STAGED = true;
    const rmi_model* m;
    uint32_t shift;
    uint8_t* s_dig;
    __device__ __forceinline__ uint32_t operator()(K k) const
    {
        return (bucket_of(*m, static_cast<uint64_t>(dev::order_bits<K, KIND, DESC>(k))) >> shift)
            & 0xFFu;
    }
};

__device__ __forceinline__ void load_model(rmi_model* s_m, const rmi_model* g_m)
{
    const uint32_t* src = reinterpret_cast<const uint32_t*>(g_m);
    uint32_t* dst = reinterpret_cast<uint32_t*>(s_m);
    for (uint32_t i = threadIdx.x; i < sizeof(rmi_model) / 4; i += blockDim.x) dst[i] = src[i];
}

//! Slope (>= 0) and intercept of y on x from moments (y = a x + c).
__device__ inline void line_from_moments(const double* mo, double& a, double& c)
{
    const double n = mo[0];
    const double mx = mo[1] / n, my = mo[2] / n;
    const double var = mo[3] / n - mx * mx;
    const double cov = mo[4] / n - mx * my;
    a = var > 0 ? fmax(cov / var, 0.0) : 0.0;
    c = my - a * mx;
}

} // namespace learned
} // namespace clc
