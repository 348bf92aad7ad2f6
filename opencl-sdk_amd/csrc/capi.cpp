// The C ABI (include/compute_c.h). Error behaviour follows the reference's
// lib/src/Compute/CApi.cpp: every call runs inside guarded() (:60-115),
// which maps each exception type to a clc_status and keeps a thread-local
// message (:49-57) cleared on success; device enumeration and context
// creation are serialized by one mutex (:117-124).
#include "../../include/compute_c.h"

#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>

#include "context.hpp"
#include "learned_sort.hpp"
#include "ops.hpp"
#include "rtc.hpp"
#include "segmented_sort.hpp"

struct clc_context_s
{
    std::unique_ptr<clc::context> ctx;
};

struct clc_graph_s
{
    std::unique_ptr<clc::seg::graph> g;
};

// A block of the device's scratch pool and its size. The pool is shared, so
// the block can outlive the context, as a reference clc_buffer outlives its
// context (CApi.cpp:21-26). The buffer is parked again on its context's
// stream while that context lives (work on the buffer was queued there), and
// with no owner stream after it: releasing a context waits for its stream.
struct clc_buffer_s
{
    std::shared_ptr<clc::stream_pool> pool;
    std::weak_ptr<const clc::stream_ref> home;
    void* data;
    std::size_t bytes;
};

namespace {

thread_local std::string last_error;
thread_local int last_hip_error = 0;

clc_status fail(clc_status status, const std::string& message, int code = 0)
{
    last_error = message;
    last_hip_error = code;
    return status;
}

template <typename F> clc_status guarded(F&& body)
{
    try
    {
        body();
        last_error.clear();
        last_hip_error = 0;
        return CLC_SUCCESS;
    } catch (const clc::no_device_error& e)
    {
        return fail(CLC_ERROR_NO_DEVICE, e.what());
    } catch (const clc::unsupported_error& e)
    {
        return fail(CLC_ERROR_UNSUPPORTED, e.what());
    } catch (const clc::build_error& e)
    {
        return fail(CLC_ERROR_BUILD, e.what());
    } catch (const clc::out_of_memory& e)
    {
        return fail(CLC_ERROR_OUT_OF_MEMORY, e.what(),
                    static_cast<int>(hipErrorOutOfMemory));
    } catch (const clc::hip_error& e)
    {
        return fail(CLC_ERROR_OPENCL, e.what(), static_cast<int>(e.code));
    } catch (const std::bad_alloc&)
    {
        return fail(CLC_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::invalid_argument& e)
    {
        return fail(CLC_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::out_of_range& e)
    {
        return fail(CLC_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::length_error& e)
    {
        return fail(CLC_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::overflow_error& e)
    {
        return fail(CLC_ERROR_OUT_OF_MEMORY, e.what());
    } catch (const std::exception& e)
    {
        return fail(CLC_ERROR_INTERNAL, e.what());
    } catch (...)
    {
        return fail(CLC_ERROR_INTERNAL, "unknown error");
    }
}

std::mutex& init_mutex()
{
    static std::mutex m;
    return m;
}

void require(bool condition, const char* message)
{
    if (!condition) throw std::invalid_argument(message);
}

// True when op and identity are forms the ahead-of-time kernels implement
// (code and id filled in). Anything else -- another expression, an
// identity that is not a constant the parser knows, a bitwise operator on
// floating-point elements -- is compiled at run time, where OpenCL C's
// rules decide, as in the reference (a bitwise float operator is a build
// error there and here).
bool canonical(int type, const char* op, const char* identity, clc::op_code& code,
               unsigned char* id)
{
    clc::type_size(type); // rejects unknown types
    try
    {
        code = clc::parse_op(op, type);
        clc::parse_identity(identity, type, id);
        return true;
    } catch (const clc::unsupported_error&)
    {
        return false;
    } catch (const std::invalid_argument&)
    {
        return false;
    }
}

std::size_t device_count_locked()
{
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e == hipErrorNoDevice || e == hipErrorInsufficientDriver
        || e == hipErrorInvalidDevice)
    {
        (void)hipGetLastError();
        return 0;
    }
    clc::hip_check(e, "hipGetDeviceCount");
    return count < 0 ? 0 : static_cast<std::size_t>(count);
}

void copy_text(char (&dst)[256], const std::string& src)
{
    const std::size_t n = src.size() < 255 ? src.size() : 255;
    std::memcpy(dst, src.data(), n);
    dst[n] = '\0';
}

void fill_info(int device, clc_device_info* info)
{
    hipDeviceProp_t prop;
    CLC_HIP(hipGetDeviceProperties(&prop, device));
    std::memset(info, 0, sizeof *info);
    copy_text(info->name, std::string(prop.name) + " (" + prop.gcnArchName + ")");
    copy_text(info->platform, "HIP");
    info->is_gpu = 1;
    info->has_fp64 = 1;
    info->global_memory = prop.totalGlobalMem;
    info->max_work_group_size = static_cast<std::size_t>(prop.maxThreadsPerBlock);
}

clc::context& get(clc_context ctx)
{
    require(ctx != nullptr && ctx->ctx != nullptr, "clc: null context");
    return *ctx->ctx;
}

std::string env_cache_dir(const char* dir)
{
    if (dir != nullptr && *dir != '\0') return dir;
    const char* env = std::getenv("CL_COMPUTE_CACHE_DIR");
    return env != nullptr ? env : "";
}

} // namespace

extern "C" {

const char* clc_last_error(void) { return last_error.c_str(); }

int clc_last_cl_error(void) { return last_hip_error; }

clc_status clc_device_count(size_t* count)
{
    if (count == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null count");
    return guarded([&] {
        std::lock_guard<std::mutex> lock(init_mutex());
        *count = device_count_locked();
    });
}

clc_status clc_device_get_info(size_t index, clc_device_info* info)
{
    if (info == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null info");
    return guarded([&] {
        std::lock_guard<std::mutex> lock(init_mutex());
        require(index < device_count_locked(), "clc: device index out of range");
        fill_info(static_cast<int>(index), info);
    });
}

clc_status clc_context_create(size_t index, clc_context* out)
{
    return clc_context_create_with_cache(index, nullptr, out);
}

clc_status clc_context_create_with_cache(size_t index, const char* cache_dir,
                                         clc_context* out)
{
    if (out == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null output");
    *out = nullptr;
    return guarded([&] {
        std::lock_guard<std::mutex> lock(init_mutex());
        const std::size_t count = device_count_locked();
        if (count == 0) throw clc::no_device_error("clc: no HIP device available");
        if (index == CLC_DEFAULT_DEVICE) index = 0; // every HIP device is a GPU
        if (index >= count)
            throw clc::no_device_error("clc: device index out of range");
        std::unique_ptr<clc_context_s> c(new clc_context_s);
        c->ctx.reset(new clc::context(static_cast<int>(index), nullptr, true,
                                      env_cache_dir(cache_dir)));
        *out = c.release();
    });
}

clc_status clc_context_create_from_stream(void* hip_stream, clc_context* out)
{
    if (out == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null output");
    *out = nullptr;
    return guarded([&] {
        std::lock_guard<std::mutex> lock(init_mutex());
        if (device_count_locked() == 0)
            throw clc::no_device_error("clc: no HIP device available");
        int device = 0;
        CLC_HIP(hipGetDevice(&device));
        std::unique_ptr<clc_context_s> c(new clc_context_s);
        c->ctx.reset(new clc::context(device, static_cast<hipStream_t>(hip_stream),
                                      false, env_cache_dir(nullptr)));
        *out = c.release();
    });
}

void clc_context_release(clc_context ctx) { delete ctx; }

clc_status clc_context_get_info(clc_context ctx, clc_device_info* info)
{
    if (ctx == nullptr || info == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null argument");
    return guarded([&] { fill_info(get(ctx).device(), info); });
}

clc_status clc_finish(clc_context ctx)
{
    if (ctx == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null context");
    return guarded([&] { CLC_HIP(hipStreamSynchronize(get(ctx).stream())); });
}

clc_status clc_context_get_stream(clc_context ctx, void** hip_stream)
{
    if (ctx == nullptr || hip_stream == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null argument");
    return guarded([&] { *hip_stream = get(ctx).stream(); });
}

clc_status clc_context_memory_stats(clc_context ctx, clc_memory_stats* stats)
{
    if (ctx == nullptr || stats == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null argument");
    return guarded([&] {
        const clc::stream_pool::figures f = get(ctx).pool()->tally();
        stats->current_bytes = f.held_bytes;
        stats->peak_bytes = f.held_peak;
        stats->total_bytes = f.obtained_bytes;
        stats->current_allocations = f.held_blocks;
        stats->total_allocations = f.obtained_blocks;
        stats->cached_bytes = f.parked_bytes;
        stats->reused = f.reused;
        stats->cross_stream = f.cross_stream;
        stats->stream_waits = f.waits;
    });
}

clc_status clc_context_program_stats(clc_context ctx, clc_program_stats* stats)
{
    return guarded([&] {
        require(ctx != nullptr && stats != nullptr, "clc: null argument");
        const clc::rtc::cache_stats& c = get(ctx).programs().stats();
        stats->memory_hits = c.memory_hits;
        stats->disk_hits = c.disk_hits;
        stats->builds = c.builds;
    });
}

clc_status clc_context_release_cache(clc_context ctx)
{
    if (ctx == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null context");
    return guarded([&] {
        clc::context& c = get(ctx);
        CLC_HIP(hipStreamSynchronize(c.stream()));
        c.pool()->trim();
    });
}

clc_status clc_context_set_profiling(clc_context ctx, int enable)
{
    if (ctx == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null context");
    return guarded([&] { get(ctx).profiler().enable(enable != 0); });
}

clc_status clc_context_kernel_timings(clc_context ctx, clc_kernel_timing* out,
                                      size_t capacity, size_t* count)
{
    if (ctx == nullptr || count == nullptr || (capacity != 0 && out == nullptr))
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null argument");
    return guarded([&] {
        const auto& totals = get(ctx).profiler().collect();
        std::size_t i = 0;
        for (const auto& kv : totals)
        {
            if (i < capacity)
            {
                clc_kernel_timing& t = out[i];
                std::memset(&t, 0, sizeof t);
                const std::size_t len = kv.first.size() < 63 ? kv.first.size() : 63;
                std::memcpy(t.name, kv.first.data(), len);
                t.launches = kv.second.launches;
                t.total_ms = kv.second.total_ms;
                t.min_ms = kv.second.min_ms;
                t.max_ms = kv.second.max_ms;
            }
            ++i;
        }
        *count = totals.size();
    });
}

clc_status clc_context_reset_timings(clc_context ctx)
{
    if (ctx == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null context");
    return guarded([&] { get(ctx).profiler().reset(); });
}

clc_status clc_reduce(clc_context ctx, const void* in, size_t n, clc_type type,
                      const char* op, const char* identity, const void* init,
                      void* result)
{
    return guarded([&] {
        require(ctx != nullptr && op != nullptr && identity != nullptr
                    && init != nullptr && result != nullptr,
                "clc: null argument");
        clc::op_code code;
        unsigned char id[8];
        if (canonical(type, op, identity, code, id))
            clc::detail::reduce(get(ctx), in, n, type, code, id, init, result);
        else
            clc::detail::reduce_expr(get(ctx), in, n, type, op, identity, init, result);
    });
}

clc_status clc_scan(clc_context ctx, const void* in, void* out, size_t n,
                    clc_type type, const char* op, const char* identity,
                    int exclusive)
{
    return guarded([&] {
        require(ctx != nullptr && op != nullptr && identity != nullptr,
                "clc: null argument");
        clc::op_code code;
        unsigned char id[8];
        if (canonical(type, op, identity, code, id))
            clc::detail::scan(get(ctx), in, out, n, type, code, id, exclusive != 0);
        else
            clc::detail::scan_expr(get(ctx), in, out, n, type, op, identity,
                                   exclusive != 0);
    });
}

clc_status clc_scan_noalloc(clc_context ctx, const void* in, void* out, size_t n,
                            clc_type type, const char* op, const char* identity,
                            int exclusive)
{
    return guarded([&] {
        require(ctx != nullptr && op != nullptr && identity != nullptr,
                "clc: null argument");
        clc::op_code code;
        unsigned char id[8];
        if (!canonical(type, op, identity, code, id))
            throw clc::unsupported_error(
                "clc_scan_noalloc: only the canonical operators and identities; "
                "compiling an expression allocates");
        clc::detail::scan_noalloc(get(ctx), in, out, n, type, code, id, exclusive != 0);
    });
}

clc_status clc_buffer_alloc(clc_context ctx, size_t bytes, clc_buffer* out)
{
    if (out == nullptr) return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null argument");
    *out = nullptr;
    if (ctx == nullptr) return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null context");
    return guarded([&] {
        clc::context& c = get(ctx);
        c.activate();
        std::unique_ptr<clc_buffer_s> b(
            new clc_buffer_s{ c.pool(), c.stream_token(), nullptr, bytes });
        if (bytes != 0) b->data = b->pool->take(bytes, c.stream());
        *out = b.release();
    });
}

void clc_buffer_release(clc_buffer buffer)
{
    if (buffer == nullptr) return;
    if (buffer->data != nullptr)
    {
        const std::shared_ptr<const clc::stream_ref> home = buffer->home.lock();
        buffer->pool->give(buffer->data, buffer->bytes, home ? *home : clc::any_stream);
    }
    delete buffer;
}

size_t clc_buffer_size(clc_buffer buffer)
{
    return buffer == nullptr ? 0 : buffer->bytes;
}

void* clc_buffer_data(clc_buffer buffer)
{
    return buffer == nullptr ? nullptr : buffer->data;
}

clc_status clc_buffer_write(clc_context ctx, clc_buffer buffer, size_t offset, size_t bytes,
                            const void* src)
{
    return guarded([&] {
        require(ctx != nullptr && buffer != nullptr, "clc: null argument");
        require(offset <= buffer->bytes && bytes <= buffer->bytes - offset,
                "clc: write out of bounds");
        require(bytes == 0 || src != nullptr, "clc: null source");
        if (bytes == 0) return;
        clc::context& c = get(ctx);
        c.activate();
        CLC_HIP(hipMemcpyAsync(static_cast<char*>(buffer->data) + offset, src, bytes,
                               hipMemcpyHostToDevice, c.stream()));
        CLC_HIP(hipStreamSynchronize(c.stream()));
    });
}

clc_status clc_buffer_read(clc_context ctx, clc_buffer buffer, size_t offset, size_t bytes,
                           void* dst)
{
    return guarded([&] {
        require(ctx != nullptr && buffer != nullptr, "clc: null argument");
        require(offset <= buffer->bytes && bytes <= buffer->bytes - offset,
                "clc: read out of bounds");
        require(bytes == 0 || dst != nullptr, "clc: null destination");
        if (bytes == 0) return;
        clc::context& c = get(ctx);
        c.activate();
        CLC_HIP(hipMemcpyAsync(dst, static_cast<const char*>(buffer->data) + offset, bytes,
                               hipMemcpyDeviceToHost, c.stream()));
        CLC_HIP(hipStreamSynchronize(c.stream()));
    });
}

clc_status clc_buffer_copy(clc_context ctx, clc_buffer src, size_t src_offset, clc_buffer dst,
                           size_t dst_offset, size_t bytes)
{
    return guarded([&] {
        require(ctx != nullptr && src != nullptr && dst != nullptr, "clc: null argument");
        require(src_offset <= src->bytes && bytes <= src->bytes - src_offset,
                "clc: copy source out of bounds");
        require(dst_offset <= dst->bytes && bytes <= dst->bytes - dst_offset,
                "clc: copy destination out of bounds");
        if (bytes == 0) return;
        clc::context& c = get(ctx);
        c.activate();
        CLC_HIP(hipMemcpyAsync(static_cast<char*>(dst->data) + dst_offset,
                               static_cast<const char*>(src->data) + src_offset, bytes,
                               hipMemcpyDeviceToDevice, c.stream()));
    });
}

clc_status clc_copy_if(clc_context ctx, const void* in, void* out, size_t n,
                       clc_type type, const char* predicate, size_t* count)
{
    return guarded([&] {
        require(ctx != nullptr && predicate != nullptr && count != nullptr,
                "clc: null argument");
        *count = clc::detail::copy_if(get(ctx), in, out, n, type, predicate);
    });
}

clc_status clc_fill(clc_context ctx, void* out, size_t n, clc_type type, const void* value)
{
    return guarded([&] {
        require(ctx != nullptr && value != nullptr, "clc: null argument");
        clc::detail::fill(get(ctx), out, n, type, value);
    });
}

clc_status clc_sequence(clc_context ctx, void* out, size_t n, clc_type type,
                        const void* start, const void* step)
{
    return guarded([&] {
        require(ctx != nullptr && start != nullptr && step != nullptr, "clc: null argument");
        clc::detail::sequence(get(ctx), out, n, type, start, step);
    });
}

clc_status clc_transform(clc_context ctx, const void* in, clc_type in_type, void* out,
                         clc_type out_type, size_t n, const char* expression)
{
    return guarded([&] {
        require(ctx != nullptr && expression != nullptr, "clc: null argument");
        clc::detail::transform(get(ctx), in, in_type, out, out_type, n, expression);
    });
}

clc_status clc_transform2(clc_context ctx, const void* in1, clc_type type1, const void* in2,
                          clc_type type2, void* out, clc_type out_type, size_t n,
                          const char* expression)
{
    return guarded([&] {
        require(ctx != nullptr && expression != nullptr, "clc: null argument");
        clc::detail::transform2(get(ctx), in1, type1, in2, type2, out, out_type, n,
                                expression);
    });
}

clc_status clc_transform_supported(clc_type in_type, int in2_type, clc_type out_type,
                                   const char* expression)
{
    return guarded([&] {
        require(expression != nullptr, "clc: null argument");
        clc::type_size(in_type);
        clc::type_size(out_type);
        if (in2_type >= 0) clc::type_size(in2_type);
        (void)clc::rtc::compile(
            clc::rtc::transform_source(in_type, in2_type < 0 ? -1 : in2_type, out_type,
                                       expression),
            "gfx950");
    });
}

clc_status clc_predicate_supported(clc_type type, const char* predicate)
{
    return guarded([&] {
        require(predicate != nullptr, "clc: null argument");
        clc::type_size(type); // rejects unknown types
        (void)clc::rtc::compile(clc::rtc::predicate_source(type, predicate), "gfx950");
    });
}

clc_status clc_sort(clc_context ctx, void* keys, clc_type key_type,
                    void* values, size_t n, int descending)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        clc::detail::radix_sort(get(ctx), keys, key_type, values, n,
                                descending != 0);
    });
}

clc_status clc_sort_tune(clc_context ctx, clc_type key_type, int with_values,
                         int* shape)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        const int s = clc::detail::sort_tune(get(ctx), key_type, with_values != 0);
        if (shape != nullptr) *shape = s;
    });
}

clc_status clc_histogram_even(clc_context ctx, const void* in, size_t n,
                              clc_type type, uint32_t* histogram,
                              size_t num_bins, double lower, double upper)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        clc::detail::histogram_even(get(ctx), in, n, type, histogram, num_bins,
                                    lower, upper);
    });
}

clc_status clc_segmented_sort(clc_context ctx, void* keys, clc_type key_type,
                              void* values, size_t num_items,
                              const uint64_t* offsets, size_t num_segments,
                              int descending)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        clc::detail::segmented_sort(get(ctx), keys, key_type, values, num_items,
                                    offsets, num_segments, descending != 0, true);
    });
}

clc_status clc_segmented_sort_tune(clc_context ctx, clc_type key_type,
                                   int with_values)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        clc::detail::segmented_sort_tune(get(ctx), key_type, with_values != 0);
    });
}

clc_status clc_segmented_sort_get_table(clc_context ctx, clc_type key_type,
                                        int with_values, int* classes)
{
    return guarded([&] {
        require(ctx != nullptr && classes != nullptr, "clc: null argument");
        const std::vector<int> t =
            clc::detail::segmented_sort_table(get(ctx), key_type, with_values != 0);
        std::memcpy(classes, t.data(), t.size() * sizeof(int));
    });
}

clc_status clc_segmented_sort_get_digit_bits(clc_context ctx, clc_type key_type,
                                             int with_values, int* bits)
{
    return guarded([&] {
        require(ctx != nullptr && bits != nullptr, "clc: null argument");
        const std::vector<int> t =
            clc::detail::segmented_sort_digit_bits(get(ctx), key_type, with_values != 0);
        std::memcpy(bits, t.data(), t.size() * sizeof(int));
    });
}

clc_status clc_segmented_sort_graph_create(clc_context ctx, void* keys,
                                           clc_type key_type, void* values,
                                           size_t num_items,
                                           const uint64_t* offsets,
                                           size_t num_segments, int descending,
                                           clc_graph* out)
{
    if (out == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null output");
    *out = nullptr;
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        std::unique_ptr<clc_graph_s> g(new clc_graph_s);
        g->g = clc::detail::segmented_sort_graph(get(ctx), keys, key_type, values,
                                                 num_items, offsets, num_segments,
                                                 descending != 0);
        *out = g.release();
    });
}

clc_status clc_graph_launch(clc_graph graph)
{
    if (graph == nullptr || graph->g == nullptr)
        return fail(CLC_ERROR_INVALID_ARGUMENT, "clc: null graph");
    return guarded([&] { graph->g->launch(); });
}

void clc_graph_release(clc_graph graph) { delete graph; }

namespace {
clc_status learned_sort_call(clc_context ctx, void* keys, clc_type key_type, void* values,
                             size_t n, int descending, bool single, clc_learned_stats* stats)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        clc::learned::stats st;
        clc::detail::learned_sort(get(ctx), keys, key_type, values, n, descending != 0, single,
                                  stats != nullptr ? &st : nullptr);
        if (stats != nullptr)
        {
            stats->used_model = st.used_model;
            stats->num_buckets = st.num_buckets;
            stats->sample_size = st.sample_size;
            stats->sample_fit_error = st.sample_fit_error;
            stats->max_bucket_overfill = st.max_bucket_overfill;
            stats->fallback_fraction = st.fallback_fraction;
            stats->root_log = st.root_log;
        }
    });
}
} // namespace

clc_status clc_learned_sort(clc_context ctx, void* keys, clc_type key_type,
                            void* values, size_t n, int descending,
                            clc_learned_stats* stats)
{
    return learned_sort_call(ctx, keys, key_type, values, n, descending, false, stats);
}

clc_status clc_learned_sort_single_scatter(clc_context ctx, void* keys, clc_type key_type,
                                           void* values, size_t n, int descending,
                                           clc_learned_stats* stats)
{
    return learned_sort_call(ctx, keys, key_type, values, n, descending, true, stats);
}

clc_status clc_learned_segmented_sort(clc_context ctx, void* keys, clc_type key_type,
                                      void* values, size_t num_items,
                                      const uint64_t* offsets, size_t num_segments,
                                      int descending, clc_learned_batched_stats* stats)
{
    return guarded([&] {
        require(ctx != nullptr, "clc: null context");
        clc::learned::batched_stats st;
        clc::detail::learned_segmented_sort(get(ctx), keys, key_type, values, num_items,
                                            offsets, num_segments, descending != 0,
                                            stats != nullptr ? &st : nullptr);
        if (stats != nullptr)
        {
            stats->buckets = st.buckets;
            stats->sample_size = st.sample_size;
            stats->fit_rejected = st.fit_rejected;
            stats->max_bucket = st.max_bucket;
            stats->fallback_fraction = st.fallback_fraction;
        }
    });
}

clc_status clc_operator_supported(clc_type type, const char* op,
                                  const char* identity)
{
    return guarded([&] {
        require(op != nullptr && identity != nullptr, "clc: null argument");
        clc::op_code code;
        unsigned char id[8];
        if (!canonical(type, op, identity, code, id))
            (void)clc::rtc::compile(
                clc::rtc::operator_source(type, op, identity,
                                          clc::rtc::WITH_SCAN | clc::rtc::WITH_REDUCE),
                "gfx950");
    });
}

} // extern "C"
