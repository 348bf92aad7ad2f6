/*
 * C interface of the MI355X (gfx950) sort / scan / reduce / histogram
 * library: a drop-in for the radix sort, scan, reduce and histogram entry
 * points of the OpenCL Compute library (reference
 * lib/include/CL/Compute/compute_c.h).
 *
 * Differences from the reference boundary, all deliberate:
 *   - The algorithms take HIP device pointers (void*) where the reference
 *     takes clc_buffer handles; the handles remain (clc_buffer_*) as
 *     library-owned memory, passed to the algorithms as clc_buffer_data().
 *     The command queue is a hipStream_t bound to the context at creation
 *     (clc_context_create_from_stream). Bounds cannot be checked on raw
 *     pointers: the caller guarantees that n elements fit.
 *   - clc_last_cl_error() returns the hipError_t of the last
 *     CLC_ERROR_OPENCL, which this library reuses for HIP runtime errors.
 *   - 64-bit sort keys (I64/U64/F64) are accepted (the reference returns
 *     CLC_ERROR_INVALID_ARGUMENT for them, CApi.cpp:522-524).
 *   - Operator and identity strings are C++ expressions over a and b of the
 *     element type T with OpenCL C's scalar builtins and limit macros; the
 *     canonical ops:: forms run ahead-of-time kernels, any other expression
 *     is compiled once per context with hipRTC (CLC_ERROR_BUILD with the
 *     build log when it does not compile).
 *   - clc_histogram_even, clc_segmented_sort and clc_learned_sort have no C
 *     entry point in the reference.
 *
 * Every function that can fail returns a clc_status; on failure
 * clc_last_error() describes it until the next call on the same thread. No
 * C++ exception crosses this boundary. Contexts are not thread-safe: use one
 * per thread. sort, scan, histogram and the segmented/learned sorts are
 * asynchronous on the context's stream; reduce blocks until its result is on
 * the host, as in the reference (Algorithm.cpp:440-441).
 */
#ifndef CLC_HIP_COMPUTE_C_H
#define CLC_HIP_COMPUTE_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clc_context_s* clc_context;

/* Status codes, numerically identical to the reference
 * (compute_c.h:30-41). */
typedef enum clc_status
{
    CLC_SUCCESS = 0,
    CLC_ERROR_OPENCL = 1, /* HIP runtime error: clc_last_cl_error() has it */
    CLC_ERROR_INVALID_ARGUMENT = 2,
    CLC_ERROR_BUILD = 3,
    CLC_ERROR_NO_DEVICE = 4,
    CLC_ERROR_OUT_OF_MEMORY = 5,
    CLC_ERROR_UNSUPPORTED = 6,
    CLC_ERROR_INTERNAL = 7,
    CLC_ERROR_IO = 8
} clc_status;

/* Element types, numerically identical to the reference (compute_c.h:44-56). */
typedef enum clc_type
{
    CLC_TYPE_I8 = 0,
    CLC_TYPE_U8 = 1,
    CLC_TYPE_I16 = 2,
    CLC_TYPE_U16 = 3,
    CLC_TYPE_I32 = 4,
    CLC_TYPE_U32 = 5,
    CLC_TYPE_I64 = 6,
    CLC_TYPE_U64 = 7,
    CLC_TYPE_F32 = 8,
    CLC_TYPE_F64 = 9
} clc_type;

/* Device index selecting the first GPU (compute_c.h:60). */
#define CLC_DEFAULT_DEVICE ((size_t)-1)

/* ---- errors (replaces compute_c.h:64-66) ------------------------------- */

const char* clc_last_error(void);
/* hipError_t of the last CLC_ERROR_OPENCL / CLC_ERROR_OUT_OF_MEMORY raised by
 * the HIP runtime, else 0. */
int clc_last_cl_error(void);

/* ---- devices and contexts (replaces compute_c.h:70-99) ------------------ */

typedef struct clc_device_info
{
    char name[256];
    char platform[256]; /* "HIP" */
    int is_gpu;
    int has_fp64;
    unsigned long long global_memory;
    size_t max_work_group_size;
} clc_device_info;

/* HIP devices. Returns CLC_SUCCESS with *count = 0 when there is none
 * (reference clc_device_count, compute_c.h:82, CApi.cpp:228-233). */
clc_status clc_device_count(size_t* count);
/* (reference clc_device_get_info, compute_c.h:83-84, CApi.cpp:235-244) */
clc_status clc_device_get_info(size_t index, clc_device_info* info);

/* Creates a context on a device: a new in-order hipStream_t, a pool memory
 * resource for temporary storage and the launch-configuration cache.
 * (reference clc_context_create, compute_c.h:89, CApi.cpp:246-249) */
clc_status clc_context_create(size_t index, clc_context* out);
/* Same; cache_dir (an existing directory, or NULL/"" for the environment
 * variable CL_COMPUTE_CACHE_DIR) persists the run-time compiled operator
 * programs of clc_scan / clc_reduce (code objects) and the autotuner's
 * launch configurations.
 * (reference clc_context_create_with_cache, compute_c.h:92-94,
 * CApi.cpp:251-283) */
clc_status clc_context_create_with_cache(size_t index, const char* cache_dir,
                                         clc_context* out);
/* Binds a caller-owned hipStream_t (NULL = the null stream) on the current
 * device. The queue-to-stream swap of compute_context(cl::CommandQueue)
 * (reference Context.hpp:33-36, MemoryResource.cpp:169-185). */
clc_status clc_context_create_from_stream(void* hip_stream, clc_context* out);
/* (reference compute_c.h:95) */
void clc_context_release(clc_context ctx);
/* (reference compute_c.h:96-97) */
clc_status clc_context_get_info(clc_context ctx, clc_device_info* info);
/* hipStreamSynchronize on the context's stream (reference compute_c.h:99). */
clc_status clc_finish(clc_context ctx);
/* The context's hipStream_t. */
clc_status clc_context_get_stream(clc_context ctx, void** hip_stream);

/* Counters of the scratch pool behind the context (the role of the
 * reference's statistics_adaptor counters, MemoryResource.hpp:123-130, and
 * pool_resource::cached_bytes, :100-101). The pool belongs to the device and
 * is shared by every context on it, so the figures cover all of them:
 * device memory held now / at most / ever obtained, in bytes and blocks;
 * the bytes parked for reuse; how many requests were served from the cache,
 * how many of those took a block another stream had parked, and how many of
 * those made the requesting stream wait on the device for the other one. */
typedef struct clc_memory_stats
{
    size_t current_bytes;
    size_t peak_bytes;
    size_t total_bytes;
    size_t current_allocations;
    size_t total_allocations;
    size_t cached_bytes;
    size_t reused;
    size_t cross_stream;
    size_t stream_waits;
} clc_memory_stats;
clc_status clc_context_memory_stats(clc_context ctx, clc_memory_stats* stats);
/* Synchronizes the context's stream, then returns every parked block of the
 * device's pool to the device (waiting for the streams that parked them;
 * the role of pool_resource::release, MemoryResource.hpp:97-98). */
clc_status clc_context_release_cache(clc_context ctx);

/* Counters of the context's cache of run-time compiled operator programs
 * (the reference's program_cache::statistics, ProgramCache.hpp:57-62):
 * programs found loaded, loaded from the cache directory, and compiled. */
typedef struct clc_program_stats
{
    size_t memory_hits;
    size_t disk_hits;
    size_t builds;
} clc_program_stats;
clc_status clc_context_program_stats(clc_context ctx, clc_program_stats* stats);

/* ---- buffers (replaces compute_c.h:100-117) ----------------------------- */
/* Library-owned device memory for callers that do not manage their own:
 * a block of the context's pool. The algorithms take device pointers, so a
 * buffer is passed to them as clc_buffer_data(buffer). The block may outlive
 * the context (it keeps the pool alive). */
typedef struct clc_buffer_s* clc_buffer;

/* bytes may be 0: the handle is valid and holds no device memory.
 * (reference clc_buffer_alloc, compute_c.h:103-104, CApi.cpp:301-312) */
clc_status clc_buffer_alloc(clc_context ctx, size_t bytes, clc_buffer* out);
/* Returns the block to the pool; NULL is a no-op. Work already queued on
 * the context's stream may still use it: the pool hands it out again only
 * to work queued after. (reference compute_c.h:105, CApi.cpp:314-320) */
void clc_buffer_release(clc_buffer buffer);
/* (reference compute_c.h:106, CApi.cpp:322-325) */
size_t clc_buffer_size(clc_buffer buffer);
/* The HIP device pointer of the buffer (NULL for a 0-byte buffer). */
void* clc_buffer_data(clc_buffer buffer);
/* Blocking transfers on the context's stream; offset and bytes must stay
 * within the buffer (CLC_ERROR_INVALID_ARGUMENT otherwise).
 * (reference compute_c.h:109-114, CApi.cpp:327-355) */
clc_status clc_buffer_write(clc_context ctx, clc_buffer buffer, size_t offset, size_t bytes,
                            const void* src);
clc_status clc_buffer_read(clc_context ctx, clc_buffer buffer, size_t offset, size_t bytes,
                           void* dst);
/* Device-to-device copy queued on the context's stream (not blocking, like
 * the reference's enqueueCopyBuffer). (reference compute_c.h:115-117,
 * CApi.cpp:357-370) */
clc_status clc_buffer_copy(clc_context ctx, clc_buffer src, size_t src_offset, clc_buffer dst,
                           size_t dst_offset, size_t bytes);

/* ---- kernel timing ------------------------------------------------------ */
/* Per-kernel hipEvent timing on the context's stream, off by default (the
 * reference's samples time kernels through OpenCL event profiling,
 * lib/include/CL/Utils/Event.hpp:14-19). */
typedef struct clc_kernel_timing
{
    char name[64];
    size_t launches;
    double total_ms;
    double min_ms;
    double max_ms;
} clc_kernel_timing;
clc_status clc_context_set_profiling(clc_context ctx, int enable);
/* Waits for the timed launches and copies up to capacity entries; *count
 * receives the number of distinct kernels timed so far. */
clc_status clc_context_kernel_timings(clc_context ctx, clc_kernel_timing* out,
                                      size_t capacity, size_t* count);
clc_status clc_context_reset_timings(clc_context ctx);

/* ---- algorithms (replaces compute_c.h:150-171) -------------------------- */
/* in/out/keys/values/histogram are device pointers; init/result point to
 * one host element of the given type. */

/* result = op(init, in[0] op ... op in[n-1]); n == 0 gives init. op must be
 * associative and commutative, identity its identity element. op and
 * identity are expressions as in the reference's binary_op
 * (Algorithm.hpp:27-31): the canonical ops:: forms run ahead-of-time
 * kernels; any other is compiled at run time (once per context, and kept in
 * the cache directory when the context has one). Blocks until the result is
 * on the host.
 * (reference clc_reduce, compute_c.h:156-159, CApi.cpp:486-499) */
clc_status clc_reduce(clc_context ctx, const void* in, size_t n, clc_type type,
                      const char* op, const char* identity, const void* init,
                      void* result);

/* Inclusive (exclusive = 0) or exclusive prefix scan; out may alias in; the
 * exclusive scan starts from identity. Associative operators, commutative or
 * not. (reference clc_scan, compute_c.h:160-163, CApi.cpp:501-515) */
clc_status clc_scan(clc_context ctx, const void* in, void* out, size_t n,
                    clc_type type, const char* op, const char* identity,
                    int exclusive);

/* clc_scan that allocates nothing, for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same arguments and
 * results as clc_scan, but only the canonical ops:: operators and
 * identities (an expression would be compiled, which allocates:
 * CLC_ERROR_UNSUPPORTED) and no profiler events. Its look-back scratch is
 * a static device array of 4 MiB + 256 bytes per GPU, held until the
 * process ends, which covers every n up to 2^31-1 for every element type
 * (larger needs are CLC_ERROR_INVALID_ARGUMENT). All calls on one GPU share
 * that array: calls that overlap on one GPU (other streams, other threads,
 * other contexts) race and must be serialised by the caller. */
clc_status clc_scan_noalloc(clc_context ctx, const void* in, void* out, size_t n,
                            clc_type type, const char* op, const char* identity,
                            int exclusive);

/* out[i] = value for i < n (value: one host element of type).
 * (reference clc_fill, compute_c.h:142-143; Algorithm.cpp:331-343) */
clc_status clc_fill(clc_context ctx, void* out, size_t n, clc_type type, const void* value);

/* out[i] = start + (T)i * step, integer arithmetic wrapping modulo 2^bits of
 * the type; floating point as a product then a sum, each rounded.
 * (reference clc_sequence, compute_c.h:144-146; Algorithm.cpp:345-359) */
clc_status clc_sequence(clc_context ctx, void* out, size_t n, clc_type type,
                        const void* start, const void* step);

/* out[i] = (T_OUT)(expression) with x = in[i] (an OpenCL C expression; the
 * types are available as T_IN and T_OUT). Compiled once per context;
 * CLC_ERROR_BUILD with the build log when it does not compile.
 * (reference clc_transform, compute_c.h:147-150; Algorithm.cpp:361-377) */
clc_status clc_transform(clc_context ctx, const void* in, clc_type in_type, void* out,
                         clc_type out_type, size_t n, const char* expression);

/* The same with x = in1[i], y = in2[i] (types T_IN1, T_IN2, T_OUT).
 * (reference clc_transform2, compute_c.h:151-155; Algorithm.cpp:379-398) */
clc_status clc_transform2(clc_context ctx, const void* in1, clc_type type1, const void* in2,
                          clc_type type2, void* out, clc_type out_type, size_t n,
                          const char* expression);

/* Copies the elements x of in[0, n) for which predicate (an OpenCL C
 * expression of x, as the reference's) is true to the front of out, in
 * order, and stores how many in *count; blocks until the count is on the
 * host. out holds n elements and may alias in. The predicate is compiled
 * once per context (CLC_ERROR_BUILD with the log when it does not compile).
 * One pass over in (the reference: flags, scan, scatter).
 * (reference clc_copy_if, compute_c.h:169-171, CApi.cpp:533-545,
 * Algorithm.cpp:529-567) */
clc_status clc_copy_if(clc_context ctx, const void* in, void* out, size_t n,
                       clc_type type, const char* predicate, size_t* count);

/* Stable in-place radix sort of n keys (I32, U32, F32, and as an extension
 * I64, U64, F64) by the reference's order (Algorithm.cpp:459-469: floats by
 * their sign-flipped bits, so -0.0 sorts before +0.0); values may be NULL or
 * hold n 32-bit elements permuted alongside. descending != 0 sorts by the
 * complemented key, keeping equal keys in input order. n <= 2^31-1.
 * (reference clc_sort, compute_c.h:166-168, CApi.cpp:517-531) */
clc_status clc_sort(clc_context ctx, void* keys, clc_type key_type,
                    void* values, size_t n, int descending);

/* Times every onesweep tile shape on 2^26 synthetic keys of key_type (with
 * or without values) and keeps the fastest for later clc_sort calls on ctx
 * with that key width; persisted like clc_segmented_sort_tune. *shape (may
 * be NULL) receives the chosen shape index. Synchronous. No counterpart in
 * the reference, whose work-group size comes from the device limit
 * (MemoryResource.cpp:187-197). */
clc_status clc_sort_tune(clc_context ctx, clc_type key_type, int with_values,
                         int* shape);

/* Zeroes num_bins uint32 counters, then counts the samples with
 * lower <= x < upper into bin (uint)((x - lower) * scale), scale =
 * bins / (upper - lower), computed in float (double for F64 samples) and
 * clamped to num_bins - 1. lower >= upper is CLC_ERROR_INVALID_ARGUMENT.
 * (reference detail::histogram_even, Algorithm.hpp:112-116,
 * Algorithm.cpp:569-615; no C entry point there) */
clc_status clc_histogram_even(clc_context ctx, const void* in, size_t n,
                              clc_type type, uint32_t* histogram,
                              size_t num_bins, double lower, double upper);

/* ---- segmented sort (no counterpart in the reference) -------------------- */
/* The reference's equivalent is one clc_sort call per segment
 * (CApi.cpp:517-531). */

/* Stable sort of every segment [offsets[i], offsets[i+1]) of keys (and the
 * 32-bit values alongside, if values != NULL), each independently, by the
 * same key order as clc_sort. offsets is a device array of num_segments + 1
 * non-decreasing uint64 offsets, offsets[num_segments] <= num_items; each
 * segment holds fewer than 2^30 keys. Segments that break this contract are
 * left untouched (never read or written out of bounds). Key types as
 * clc_sort. Asynchronous; a fixed launch sequence with no host round trip. */
clc_status clc_segmented_sort(clc_context ctx, void* keys, clc_type key_type,
                              void* values, size_t num_items,
                              const uint64_t* offsets, size_t num_segments,
                              int descending);

/* Times every workgroup configuration for each segment-size bucket
 * (bucket b: 2^(b-1) < size <= 2^b) on synthetic segments and keeps the
 * fastest, then for an LDS configuration the widest digit of its passes
 * (8, 7, 6, 5 or 4 bits), for this key type with or without values. The
 * result is used by
 * later calls on ctx and, when the context has a cache directory, written
 * there (one file per device, key type and values flag; entries from
 * another device or runtime are ignored and re-tuned). Synchronous. */
clc_status clc_segmented_sort_tune(clc_context ctx, clc_type key_type,
                                   int with_values);

/* The size-class table in use: classes[b] for the 32 buckets, values
 * 0..6 = LDS configurations (64x4, 64x8, 128x8, 256x8, 256x16, 512x16,
 * 1024x16 threads x keys), 7 = multi-workgroup onesweep. */
clc_status clc_segmented_sort_get_table(clc_context ctx, clc_type key_type,
                                        int with_values, int* classes);

/* The digit widths in use: bits[b] for the 32 buckets, 1..8, the widest
 * digit of the LDS passes for segments of bucket b (keys that differ in
 * their low k bits take ceil(k / bits[b]) passes, the k bits split evenly
 * over them). No counterpart in the reference, which has no segmented sort. */
clc_status clc_segmented_sort_get_digit_bits(clc_context ctx, clc_type key_type,
                                             int with_values, int* bits);

/* A segmented sort of fixed buffers captured once into a hipGraph, with its
 * own scratch; each clc_graph_launch replays the whole sort as one graph
 * launch on the context's stream. The buffers' contents (keys, values and
 * offsets) may change between launches; their addresses, num_items and
 * num_segments may not. The context must outlive the graph. */
typedef struct clc_graph_s* clc_graph;
clc_status clc_segmented_sort_graph_create(clc_context ctx, void* keys,
                                           clc_type key_type, void* values,
                                           size_t num_items,
                                           const uint64_t* offsets,
                                           size_t num_segments, int descending,
                                           clc_graph* out);
clc_status clc_graph_launch(clc_graph graph);
void clc_graph_release(clc_graph graph);

/* ---- learned-bucket sort (no counterpart in the reference) ---------------- */

/* Model quality of one clc_learned_sort call. */
typedef struct clc_learned_stats
{
    int used_model; /* 0: the whole input went to the plain radix sort */
    size_t num_buckets;
    size_t sample_size;
    /* estimated fraction of keys in buckets too large for LDS holding more
     * than one distinct key (above 0.25 the model is not used) */
    double sample_fit_error;
    /* largest bucket / average bucket */
    double max_bucket_overfill;
    /* fraction of keys in buckets too large for LDS, finished by the
     * onesweep fallback instead */
    double fallback_fraction;
    /* 1 when the model's root is linear in the log image of the key (the
     * bits of the key converted to double, read as an integer) rather than
     * in the key: of the two candidate roots, the one whose sample predicts
     * fewer keys in overfull buckets is used */
    int root_log;
} clc_learned_stats;

/* Same result as clc_sort (stable, same key order, values permuted
 * alongside), computed through a learned CDF model: a two-level RMI fitted
 * to a strided sample (every 100th key up to 65,536 samples, every 400th
 * beyond; sorted by the plain radix sort) maps keys
 * monotonically to up to 65536 buckets, two stable passes keyed by the
 * bucket id group the keys, and each bucket is finished in LDS; buckets too
 * large for LDS, and inputs whose sample fits badly, go to the plain radix
 * sort. n <= 2^31-1 (inputs below 2^16 or from 2^30 keys use the plain sort
 * directly). Synchronizes once to check the fit, and once more to fill
 * stats when stats != NULL. */
clc_status clc_learned_sort(clc_context ctx, void* keys, clc_type key_type,
                            void* values, size_t n, int descending,
                            clc_learned_stats* stats);

/* clc_learned_sort with the keys grouped by one fused scatter into at most
 * 4096 buckets instead of two 8-bit passes: each workgroup counts the
 * buckets of its contiguous chunk, one exclusive scan over the counts
 * (bucket-major) fixes every bucket's size and each chunk's place in it,
 * and each workgroup then writes its chunk's keys to their buckets tile by
 * tile (a stable sort of the tile by bucket in LDS first). Same result;
 * meant for inputs up to about 2^24 keys, where 4096 buckets fit LDS (a
 * larger input leaves buckets above the LDS capacity to the onesweep
 * fallback). */
clc_status clc_learned_sort_single_scatter(clc_context ctx, void* keys, clc_type key_type,
                                           void* values, size_t n, int descending,
                                           clc_learned_stats* stats);

/* Outcome of one clc_learned_segmented_sort call. */
typedef struct clc_learned_batched_stats
{
    /* buckets of the segments sorted through a model, plus every other
     * segment (sorted whole) */
    size_t buckets;
    /* sampled keys over the segments that were modelled */
    size_t sample_size;
    /* segments whose model failed the fit check (sorted whole instead) */
    size_t fit_rejected;
    /* keys in the largest bucket or whole segment */
    size_t max_bucket;
    /* fraction of keys in buckets or segments too large for LDS, finished by
     * the segmented sort's onesweep path */
    double fallback_fraction;
} clc_learned_batched_stats;

/* Same result as clc_segmented_sort (each segment [offsets[i],
 * offsets[i+1]) sorted stably, values permuted alongside) for 32-bit keys
 * (U32, I32, F32), through one learned model per segment: segments larger
 * than the LDS segment sort holds are sampled (1%), their samples sorted by
 * the segmented sort, one RMI per segment fitted in one launch, keys
 * grouped into ~3072-key buckets by one or two stable passes (inference
 * fused in), and the buckets finished by the segmented sort; segments that
 * fit LDS, or whose sample fits badly, are sorted whole. No host round trip
 * (one synchronisation to fill stats when stats != NULL). */
clc_status clc_learned_segmented_sort(clc_context ctx, void* keys, clc_type key_type,
                                      void* values, size_t num_items,
                                      const uint64_t* offsets, size_t num_segments,
                                      int descending, clc_learned_batched_stats* stats);

/* Whether clc_scan / clc_reduce accept op/identity for type: CLC_SUCCESS
 * (a canonical form, or an expression that compiles for gfx950),
 * CLC_ERROR_BUILD (it does not compile; clc_last_error() has the build log)
 * or CLC_ERROR_INVALID_ARGUMENT. Needs no device. */
clc_status clc_operator_supported(clc_type type, const char* op,
                                  const char* identity);

/* Whether clc_transform (in2_type < 0) or clc_transform2 accepts expression
 * for the types: CLC_SUCCESS when it compiles for gfx950, CLC_ERROR_BUILD
 * with the log otherwise, or CLC_ERROR_INVALID_ARGUMENT. Needs no device. */
clc_status clc_transform_supported(clc_type in_type, int in2_type, clc_type out_type,
                                   const char* expression);

/* Whether clc_copy_if accepts predicate for type: CLC_SUCCESS when it
 * compiles for gfx950, CLC_ERROR_BUILD with the log otherwise, or
 * CLC_ERROR_INVALID_ARGUMENT. Needs no device. */
clc_status clc_predicate_supported(clc_type type, const char* predicate);

#ifdef __cplusplus
}
#endif

#endif
