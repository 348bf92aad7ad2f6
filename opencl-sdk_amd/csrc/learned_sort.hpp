// Learned-bucket sort entry point (see learned_sort.hip for the design).
#pragma once

#include "context.hpp"

namespace clc {
namespace learned {

//! Inputs below this size go to the plain radix sort directly.
constexpr std::size_t MIN_KEYS = std::size_t(1) << 16;
//! Estimated fraction of keys in overfull buckets with more than one
//! distinct key above which the model is not used.
constexpr double FIT_ERROR_LIMIT = 0.25;

//! Model quality (mirrors clc_learned_stats).
struct stats
{
    int used_model = 0;
    std::size_t num_buckets = 0;
    std::size_t sample_size = 0;
    double sample_fit_error = 0;
    double max_bucket_overfill = 0;
    double fallback_fraction = 0;
    int root_log = 0; //!< 1 when the root is linear in the key's log image
};

//! Outcome of one batched learned sort (mirrors clc_learned_batched_stats).
struct batched_stats
{
    std::size_t buckets = 0; //!< buckets of fitted segments plus segments sorted whole
    std::size_t sample_size = 0; //!< sampled keys over all fitted segments
    std::size_t fit_rejected = 0; //!< segments whose fit check failed (sorted whole)
    std::size_t max_bucket = 0; //!< keys in the largest bucket or whole segment
    double fallback_fraction = 0; //!< keys in buckets or segments too large for LDS
};

} // namespace learned

namespace detail {

//! Stable sort (same result as radix_sort) through the learned buckets;
//! fills out when given (which costs one extra synchronisation). With
//! single_scatter, keys reach at most 4096 buckets in one fused scatter
//! instead of two 8-bit passes into up to 65536.
void learned_sort(context& ctx, void* keys, int key_type, void* values, std::size_t n,
                  bool descending, bool single_scatter, learned::stats* out);

//! Stable sort of every segment [offsets[i], offsets[i+1]) of 32-bit keys
//! (and values) with one learned model per segment (learned_batched.hip);
//! same result as segmented_sort. Fills out when given (one extra
//! synchronisation).
void learned_segmented_sort(context& ctx, void* keys, int key_type, void* values,
                            std::size_t num_items, const uint64_t* offsets,
                            std::size_t num_segments, bool descending,
                            learned::batched_stats* out);

} // namespace detail
} // namespace clc
