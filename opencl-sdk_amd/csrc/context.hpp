// Execution context and launch layer: one HIP device, one in-order stream,
// the device's shared scratch pool and the device facts the launch
// configurations depend on. Plays the role of the reference's compute_context
// (lib/include/CL/Compute/Context.hpp:25-59, MemoryResource.cpp:169-197)
// with the command queue replaced by a hipStream_t.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "stream_pool.hpp"

namespace clc {

namespace rtc {
class program_cache;
}

//! A failed HIP runtime call (CLC_ERROR_OPENCL at the C boundary).
struct hip_error : std::runtime_error
{
    hip_error(hipError_t code, const std::string& what)
        : std::runtime_error(what), code(code)
    {}
    hipError_t code;
};

struct no_device_error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

//! Throws hip_error (or out_of_memory) when e is not hipSuccess.
void hip_check(hipError_t e, const char* what);

#define CLC_HIP(call) ::clc::hip_check((call), #call)

//! The scratch pool of a device, shared by every context on it and by the
//! buffers they hand out; created with the first context on the device and
//! destroyed (its blocks returned) when the last holder lets go. Its cache
//! cap is an eighth of the device's memory, at least 8 GiB.
std::shared_ptr<stream_pool> device_pool(int device, std::size_t global_memory);

//! Per-kernel launch timing with hipEvent pairs on the context's stream
//! (the reference's samples time kernels with event profiling,
//! lib/include/CL/Utils/Event.hpp:14-19). Off by default; when on, each
//! timed launch is bracketed by two events, and collect() waits for them.
class profiler {
public:
    struct stats
    {
        std::size_t launches = 0;
        double total_ms = 0.0, min_ms = 0.0, max_ms = 0.0;
    };
    ~profiler();
    bool enabled() const { return enabled_; }
    void enable(bool on) { enabled_ = on; }
    //! Records the start event; returns a token for stop(), or -1 when off.
    int start(const char* name, hipStream_t stream);
    void stop(int token, hipStream_t stream);
    //! Waits for outstanding events and folds them into the totals.
    const std::map<std::string, stats>& collect();
    void reset();

private:
    struct pending
    {
        std::string name;
        hipEvent_t begin, end;
    };
    hipEvent_t get_event();
    bool enabled_ = false;
    std::vector<pending> pending_;
    std::vector<hipEvent_t> free_;
    std::map<std::string, stats> totals_;
};

struct device_facts
{
    std::string name;
    int cu_count = 256;
    int xcd_count = 8;
    std::size_t lds_per_block = 65536;
    std::size_t global_memory = 0;
    int max_threads_per_block = 1024;
    std::string arch;
};

class context {
public:
    //! Creates a stream on device (owned) or binds stream (not owned).
    context(int device, hipStream_t stream, bool owns_stream,
            std::string cache_dir);
    ~context();
    context(const context&) = delete;
    context& operator=(const context&) = delete;

    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    //! This context's stream for holders that may outlive the context
    //! (buffers): expires when the context is destroyed.
    std::weak_ptr<const stream_ref> stream_token() const { return stream_token_; }
    const device_facts& facts() const { return facts_; }
    //! Where this context's launches take scratch from: the device pool,
    //! or a graph's own block while a graph is captured.
    const std::shared_ptr<scratch_source>& resource() const { return resource_; }
    const std::shared_ptr<stream_pool>& pool() const { return pool_; }
    const std::string& cache_dir() const { return cache_dir_; }

    //! Pinned host staging for small blocking reads (reduce results).
    void* pinned() const { return pinned_; }

    //! Makes this context's device current for the calling thread.
    void activate() const;

    class profiler& profiler() { return profiler_; }

    //! Autotuned launch tables, keyed by the tuner (key type, values, ...).
    std::map<int, std::vector<int>>& tuned() { return tuned_; }

    //! Run-time compiled operator programs (created on first use).
    rtc::program_cache& programs();

    //! Auxiliary streams (default priority, non-blocking) for independent
    //! branches of one operation, forked from and joined back to stream()
    //! with the events below; created on first use, owned by the context.
    static constexpr int AUX_STREAMS = 4;
    static constexpr int AUX_EVENTS = 16;
    hipStream_t aux_stream(int i);
    hipEvent_t aux_event(int i);

    //! Temporarily redirects launches and scratch (used while a graph is
    //! captured on a private stream with graph-owned scratch); returns the
    //! previous values so the caller can restore them.
    hipStream_t swap_stream(hipStream_t s)
    {
        hipStream_t old = stream_;
        stream_ = s;
        return old;
    }
    std::shared_ptr<scratch_source> swap_resource(std::shared_ptr<scratch_source> r)
    {
        std::shared_ptr<scratch_source> old = resource_;
        resource_ = std::move(r);
        return old;
    }

private:
    int device_;
    hipStream_t stream_;
    bool owns_stream_;
    std::string cache_dir_;
    device_facts facts_;
    std::shared_ptr<stream_pool> pool_;
    std::shared_ptr<scratch_source> resource_;
    std::shared_ptr<const stream_ref> stream_token_;
    void* pinned_ = nullptr;
    class profiler profiler_;
    std::map<int, std::vector<int>> tuned_;
    std::unique_ptr<rtc::program_cache> programs_;
    hipStream_t aux_streams_[AUX_STREAMS] = {};
    hipEvent_t aux_events_[AUX_EVENTS] = {};
};

//! Brackets one kernel launch with profiler events when profiling is on.
class timed_launch {
public:
    timed_launch(context& ctx, const char* name)
        : ctx_(ctx), token_(ctx.profiler().start(name, ctx.stream()))
    {}
    ~timed_launch()
    {
        if (token_ >= 0) ctx_.profiler().stop(token_, ctx_.stream());
    }
    timed_launch(const timed_launch&) = delete;
    timed_launch& operator=(const timed_launch&) = delete;

private:
    context& ctx_;
    int token_;
};

// ---- on-disk cache of tuned launch tables ----------------------------------
// One file per table name in the context's cache directory, in the format of
// tuning_file.hpp (identity: device architecture, CU count, HIP runtime
// version); a missing, foreign or damaged file is ignored (and re-tuned).

//! Reads table name; false when there is no cache directory, no file, or a
//! stale/foreign one.
bool load_tuning(const context& ctx, const std::string& name, std::vector<int>& out);
//! Writes table name; failures are ignored (the cache is an optimisation).
void store_tuning(const context& ctx, const std::string& name,
                  const std::vector<int>& values);

//! Kernels index sizes up to 2^31-1 per call, as the reference
//! (lib/src/Compute/Detail.hpp:18-23).
inline void check_size(std::size_t n)
{
    if (n > 0x7FFFFFFFu)
        throw std::length_error("clc: at most 2^31-1 elements per call");
}

namespace detail {

// Entry points behind the C ABI (the reference's detail:: functions,
// Algorithm.hpp:96-116). Pointers are device pointers.
void radix_sort(context& ctx, void* keys, int key_type, void* values,
                std::size_t n, bool descending);
void scan(context& ctx, const void* in, void* out, std::size_t n, int type,
          int op, const void* identity, bool exclusive);
//! scan with its look-back scratch in a static per-GPU device array instead
//! of the pool: no device allocation on any call. Canonical operators only.
//! Calls that overlap on one GPU race for the array.
void scan_noalloc(context& ctx, const void* in, void* out, std::size_t n, int type,
                  int op, const void* identity, bool exclusive);
//! Looks up the address of scan_noalloc's array on device (which must be
//! current); once per device, at context creation.
hipError_t scan_noalloc_setup(int device);
void reduce(context& ctx, const void* in, std::size_t n, int type, int op,
            const void* identity, const void* init, void* result);
//! scan / reduce with an operator and identity given as OpenCL C
//! expressions, compiled at run time (the context's program cache).
void scan_expr(context& ctx, const void* in, void* out, std::size_t n, int type,
               const char* op, const char* identity, bool exclusive);
void reduce_expr(context& ctx, const void* in, std::size_t n, int type,
                 const char* op, const char* identity, const void* init,
                 void* result);
//! Copies the elements x of in[0, n) with (predicate) true, an OpenCL C
//! expression of x, to the front of out in order; returns how many (blocks
//! until the count is on the host). out may alias in.
std::size_t copy_if(context& ctx, const void* in, void* out, std::size_t n, int type,
                    const char* predicate);
//! Elementwise helpers (elementwise.hip): value / start / step point to one
//! host element of type; expressions are OpenCL C over x (and y).
void fill(context& ctx, void* out, std::size_t n, int type, const void* value);
void sequence(context& ctx, void* out, std::size_t n, int type, const void* start,
              const void* step);
void transform(context& ctx, const void* in, int in_type, void* out, int out_type,
               std::size_t n, const char* expression);
void transform2(context& ctx, const void* in1, int type1, const void* in2, int type2,
                void* out, int out_type, std::size_t n, const char* expression);
void histogram_even(context& ctx, const void* in, std::size_t n, int type,
                    uint32_t* histogram, std::size_t num_bins, double lower,
                    double upper);
//! Zeroes bytes of device memory with a kernel (graph-capturable, unlike a
//! memset node).
void zero_fill(hipStream_t stream, void* p, std::size_t bytes);
//! Times every onesweep tile shape on 2^26 synthetic keys of key_type and
//! keeps the fastest for this key width and values flag; returns it.
int sort_tune(context& ctx, int key_type, bool with_values);
int sort_tuning_slot(std::size_t kbytes, bool values);
std::string sort_tuning_name(std::size_t kbytes, bool values);

} // namespace detail

} // namespace clc
