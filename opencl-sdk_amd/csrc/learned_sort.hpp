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

} // namespace learned

namespace detail {

//! Stable sort (same result as radix_sort) through the learned buckets;
//! fills out when given (which costs one extra synchronisation).
void learned_sort(context& ctx, void* keys, int key_type, void* values, std::size_t n,
                  bool descending, learned::stats* out);

} // namespace detail
} // namespace clc
