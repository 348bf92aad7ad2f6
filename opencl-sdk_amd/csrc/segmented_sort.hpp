// Segmented (batched) stable radix sort for gfx950: thousands of uneven
// segments sorted in one fixed-shape launch sequence that needs no host
// round trip and can be captured into a hipGraph.
//
//   classify      each segment goes to a size class by the tuned table
//                 (bucket = ceil(log2 m)): one of the LDS configurations
//                 below, or "large"; per-class lists are filled by one
//                 atomic per class and wave
//   lds classes   one persistent kernel per configuration; each workgroup
//                 takes segments from its class list and sorts each one
//                 wholly in LDS (lds_sort_segment): one read and one write
//                 of HBM per key
//   large         segments above every LDS capacity: a single-block scan
//                 lays out their onesweep tiles, one persistent kernel
//                 counts every digit place per segment, one scans the
//                 counts, a single-block scan keeps the segments with some
//                 varying place (one repeated key needs nothing), and one
//                 persistent onesweep kernel per digit place processes their
//                 tiles, with look-back chains per segment
//
// The reference library has no segmented sort; its per-segment equivalent
// is one detail::radix_sort call per segment (Algorithm.cpp:453-527).
#pragma once

#include "context.hpp"
#include "device_common.hpp"
#include "radix_tile.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace clc {
namespace seg {

//! LDS configurations (workgroup size, items per thread); a segment of m
//! keys fits configuration c when m <= BLOCK * ITEMS.
constexpr int NCFG = 7;
constexpr int CFG_BLOCK[NCFG] = { 64, 64, 128, 256, 256, 512, 1024 };
constexpr int CFG_ITEMS[NCFG] = { 4, 8, 8, 8, 16, 16, 16 };
//! Class index of the large (onesweep) path.
constexpr int LARGE = NCFG;
//! An extra LDS shape for segments just above 4096 keys: bucket 13 holds
//! 4097..8192, and a configuration that holds 8192 (512 x 16) runs a
//! segment of 4097..6144 half empty (2^28 keys in 4097-key segments took
//! 1.59x the time of 4096-key ones). When bucket 13's class is an LDS
//! configuration, its segments of at most MID_CAP keys take this shape
//! instead. Not a table entry (the tuned classes keep their meaning); class
//! index MID internally.
constexpr int MID = NCFG + 1;
constexpr int MID_BLOCK = 256;
constexpr int MID_ITEMS = 24;
constexpr uint32_t MID_CAP = MID_BLOCK * MID_ITEMS;
constexpr int MID_BUCKET = 13;
//! Segment-size buckets: bucket b holds 2^(b-1) < m <= 2^b (m >= 2).
constexpr int NBUCKET = 32;
//! The sort's work counters sit CTR_STRIDE words (256 bytes) apart, one
//! per slot, CTR_SLOTS slots: device-scope atomics on neighbouring words of
//! one line are serialised with each other.
constexpr uint32_t CTR_STRIDE = 64;
constexpr int CTR_SLOTS = 64;
//! Onesweep tile of the large path.
constexpr int LARGE_BLOCK = 1024;
//! Keys per thread of a large-path onesweep tile: as many as the persistent
//! kernel keeps in 128 VGPRs without spilling (32-bit keys: 16; 18 and 20
//! spill). 65,536 lognormal segments: 1.058 ms at 16 against 1.067 at 12 and
//! 1.106 at 8; the batched float segments' plain sort 34.8 against 39.1-39.8
//! ms at 12.
constexpr int large_items(std::size_t kbytes) { return kbytes == 8 ? 8 : 16; }
constexpr uint32_t large_tile(std::size_t kbytes)
{
    return static_cast<uint32_t>(LARGE_BLOCK * large_items(kbytes));
}

//! Dynamic LDS bytes of the large path's onesweep tile: per-wave digit
//! counters, digit bases, and one staging buffer that the keys and then the
//! values pass through (split staging).
inline std::size_t large_lds(std::size_t kbytes)
{
    return std::size_t(LARGE_BLOCK / 64) * 1024 + 256 * 4 + large_tile(kbytes) * kbytes + 64;
}

inline std::size_t cfg_capacity(int c)
{
    return std::size_t(CFG_BLOCK[c]) * CFG_ITEMS[c];
}

//! Dynamic LDS bytes of an LDS shape (block, capacity) for key size kbytes.
constexpr std::size_t shape_lds(int block, std::size_t tile, std::size_t kbytes, bool values)
{
    return std::size_t(block / 64) * 1024
        + tile * radix::lds_segment_bytes_per_key(kbytes, values) + 64;
}

//! Dynamic LDS bytes of configuration c for key size kbytes.
inline std::size_t cfg_lds(int c, std::size_t kbytes, bool values)
{
    return shape_lds(CFG_BLOCK[c], cfg_capacity(c), kbytes, values);
}

inline std::size_t mid_lds(std::size_t kbytes, bool values)
{
    return shape_lds(MID_BLOCK, MID_CAP, kbytes, values);
}

//! Class per bucket; the autotuner fills it, defaults come from
//! default_table().
//! Per bucket: its size class, and the widest digit (1..8 bits) of the
//! LDS passes for its segments (a segment whose keys differ in `bits` low
//! bits takes ceil(bits / digit_bits[b]) passes).
struct table
{
    int cls[NBUCKET];
    int digit_bits[NBUCKET];
};
constexpr int MAX_DIGIT_BITS = 8;

table default_table(std::size_t kbytes, bool values, std::size_t lds_limit);

//! Enqueues the segmented sort for one key type on ctx's stream, or with
//! prepare set only sizes its scratch and warms its launch state. Returns
//! the scratch bytes. Instantiated in segmented_k32.hip / segmented_k64.hip.
//! With overlap_large the onesweep passes of segments above every LDS
//! capacity run while the LDS classes are still sorting (on the 65,536
//! lognormal segments of the batched benchmark: 1.47 -> 1.38 ms); without
//! it they follow the classes, which the learned sorts' bucket finish
//! prefers (2^28 lognormal / Gaussian-mixture keys: 6.44 / 8.0 ms against
//! 6.58 / 8.4 ms overlapped: their few overfull buckets' chains stall
//! behind the classes' workgroups).
template <typename K, int KIND>
std::size_t sort_typed(context& ctx, bool prepare, void* keys, bool with_values,
                       void* values, std::size_t num_items, const uint64_t* offsets,
                       std::size_t num_segments, bool descending, const table& t,
                       bool overlap_large);

//! A captured segmented sort: the whole launch sequence as one hipGraph,
//! replayed on the context's stream. Its scratch is one block of the device
//! pool held for the graph's lifetime and parked again on the context's
//! stream (where the replays were queued) when the graph is destroyed.
struct graph
{
    context* ctx = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    std::unique_ptr<temp_buffer> scratch;
    hipStream_t capture_stream = nullptr;
    ~graph();
    void launch();
};

} // namespace seg

namespace detail {

//! Stable sort of every segment [offsets[i], offsets[i+1]) of keys/values.
//! overlap_large: see seg::sort_typed (the C entry point sets it; the
//! learned sorts' bucket finish does not).
void segmented_sort(context& ctx, void* keys, int key_type, void* values,
                    std::size_t num_items, const uint64_t* offsets,
                    std::size_t num_segments, bool descending, bool overlap_large = false);
//! Runs the autotuner for a key type and stores its table in ctx (and on
//! disk when ctx has a cache directory).
void segmented_sort_tune(context& ctx, int key_type, bool with_values);
//! The size-class table the segmented sort will use.
std::vector<int> segmented_sort_table(context& ctx, int key_type, bool with_values);
//! The widest digit (bits) of the LDS passes per segment-size bucket.
std::vector<int> segmented_sort_digit_bits(context& ctx, int key_type, bool with_values);
//! Captures the segmented sort of fixed buffers into a graph.
std::unique_ptr<seg::graph> segmented_sort_graph(context& ctx, void* keys, int key_type,
                                                 void* values, std::size_t num_items,
                                                 const uint64_t* offsets,
                                                 std::size_t num_segments, bool descending);

} // namespace detail
} // namespace clc
