#include "context.hpp"
#include "rtc.hpp"
#include "tuning_file.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace clc {

void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    (void)hipGetLastError(); // clear the sticky per-thread error state
    const std::string msg =
        std::string(what) + " failed: " + hipGetErrorString(e) + " ("
        + std::to_string(static_cast<int>(e)) + ")";
    if (e == hipErrorOutOfMemory) throw out_of_memory(msg);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
        throw no_device_error(msg);
    throw hip_error(e, msg);
}

namespace {

//! The pool's view of one HIP device.
class hip_link final : public device_link {
public:
    explicit hip_link(int device) : device_(device) {}
    ~hip_link() override
    {
        if (mark_) (void)hipEventDestroy(mark_);
    }

    void* obtain(std::size_t bytes) override
    {
        on_device here(device_);
        void* p = nullptr;
        hip_check(hipMalloc(&p, bytes == 0 ? 1 : bytes), "hipMalloc");
        return p;
    }
    void surrender(void* p, std::size_t) noexcept override { (void)hipFree(p); }
    bool idle(stream_ref s) override
    {
        const hipError_t e = hipStreamQuery(static_cast<hipStream_t>(s));
        if (e == hipSuccess) return true;
        (void)hipGetLastError(); // hipErrorNotReady is not an error here
        return false;
    }
    void follow(stream_ref later, stream_ref earlier) override
    {
        // The pool serialises its calls, so one event serves every follow:
        // a stream wait binds to the record that precedes it, and a later
        // record of the same event does not move it.
        if (!mark_)
        {
            on_device here(device_);
            CLC_HIP(hipEventCreateWithFlags(&mark_, hipEventDisableTiming));
        }
        CLC_HIP(hipEventRecord(mark_, static_cast<hipStream_t>(earlier)));
        CLC_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(later), mark_, 0));
    }
    void drain(stream_ref s) noexcept override
    {
        (void)hipStreamSynchronize(static_cast<hipStream_t>(s));
    }

private:
    //! Makes device current for a scope and restores the caller's device.
    struct on_device
    {
        explicit on_device(int d)
        {
            (void)hipGetDevice(&prev);
            if (prev != d) (void)hipSetDevice(d);
            dev = d;
        }
        ~on_device()
        {
            if (prev != dev) (void)hipSetDevice(prev);
        }
        int prev = -1, dev = -1;
    };
    int device_;
    hipEvent_t mark_ = nullptr;
};

} // namespace

std::shared_ptr<stream_pool> device_pool(int device, std::size_t global_memory)
{
    static std::mutex m;
    static std::map<int, std::weak_ptr<stream_pool>> pools;
    std::lock_guard<std::mutex> lock(m);
    if (auto live = pools[device].lock()) return live;
    // Cache up to an eighth of device memory (36 GB on a 288 GB MI355X): a
    // 2^28 u64 key-value sort and a learned sort each keep 3+ GiB of
    // alternate buffers, and a workload that alternates between them (or
    // follows a u32 sort) must not return a block to hipFree and take it
    // back from hipMalloc on every call (0.5 ms per 2^28-key learned sort).
    stream_pool::limits l;
    l.cache_cap = std::max(l.cache_cap, global_memory / 8);
    auto fresh = std::make_shared<stream_pool>(std::make_shared<hip_link>(device), l);
    pools[device] = fresh;
    return fresh;
}

context::context(int device, hipStream_t stream, bool owns_stream,
                 std::string cache_dir)
    : device_(device), stream_(stream), owns_stream_(owns_stream),
      cache_dir_(std::move(cache_dir))
{
    CLC_HIP(hipSetDevice(device_));
    hipDeviceProp_t prop;
    CLC_HIP(hipGetDeviceProperties(&prop, device_));
    facts_.name = prop.name;
    facts_.arch = prop.gcnArchName;
    facts_.cu_count = prop.multiProcessorCount;
    facts_.lds_per_block = prop.sharedMemPerBlock;
    facts_.global_memory = prop.totalGlobalMem;
    facts_.max_threads_per_block = prop.maxThreadsPerBlock;
    facts_.xcd_count =
        facts_.arch.compare(0, 5, "gfx95") == 0
                || facts_.arch.compare(0, 5, "gfx94") == 0
            ? 8
            : 1;
    if (owns_stream_) CLC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    pool_ = device_pool(device_, facts_.global_memory);
    resource_ = pool_;
    stream_token_ = std::make_shared<const stream_ref>(static_cast<stream_ref>(stream_));
    CLC_HIP(hipHostMalloc(&pinned_, 64, hipHostMallocDefault));
    // The allocation-free scan's array address on this device. A failure
    // here is reported by that entry point, which retries the lookup.
    (void)detail::scan_noalloc_setup(device_);
}

context::~context()
{
    (void)hipSetDevice(device_);
    stream_token_.reset();
    (void)hipStreamSynchronize(stream_);
    // The pool outlives this context (other contexts and buffers share it):
    // blocks parked by this context's streams must not name them any more.
    pool_->retire(stream_);
    for (hipStream_t& a : aux_streams_)
        if (a)
        {
            pool_->retire(a);
            (void)hipStreamSynchronize(a);
            (void)hipStreamDestroy(a);
        }
    for (hipEvent_t& e : aux_events_)
        if (e) (void)hipEventDestroy(e);
    programs_.reset(); // unloads run-time compiled modules
    if (pinned_) (void)hipHostFree(pinned_);
    resource_.reset();
    pool_.reset(); // the last holder returns the parked blocks
    if (owns_stream_) (void)hipStreamDestroy(stream_);
}

profiler::~profiler()
{
    for (auto& p : pending_)
    {
        (void)hipEventDestroy(p.begin);
        (void)hipEventDestroy(p.end);
    }
    for (hipEvent_t e : free_) (void)hipEventDestroy(e);
}

hipEvent_t profiler::get_event()
{
    if (!free_.empty())
    {
        hipEvent_t e = free_.back();
        free_.pop_back();
        return e;
    }
    hipEvent_t e;
    CLC_HIP(hipEventCreate(&e));
    return e;
}

int profiler::start(const char* name, hipStream_t stream)
{
    if (!enabled_) return -1;
    pending p{ name, get_event(), get_event() };
    CLC_HIP(hipEventRecord(p.begin, stream));
    pending_.push_back(p);
    return static_cast<int>(pending_.size() - 1);
}

void profiler::stop(int token, hipStream_t stream)
{
    if (token < 0 || token >= static_cast<int>(pending_.size())) return;
    (void)hipEventRecord(pending_[token].end, stream);
}

const std::map<std::string, profiler::stats>& profiler::collect()
{
    for (auto& p : pending_)
    {
        CLC_HIP(hipEventSynchronize(p.end));
        float ms = 0.0f;
        CLC_HIP(hipEventElapsedTime(&ms, p.begin, p.end));
        stats& s = totals_[p.name];
        s.min_ms = s.launches == 0 ? ms : std::min(s.min_ms, double(ms));
        s.max_ms = s.launches == 0 ? ms : std::max(s.max_ms, double(ms));
        s.total_ms += ms;
        ++s.launches;
        free_.push_back(p.begin);
        free_.push_back(p.end);
    }
    pending_.clear();
    return totals_;
}

void profiler::reset()
{
    collect();
    totals_.clear();
}

namespace {

std::string tuning_identity(const context& ctx)
{
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    return ctx.facts().arch + " cu" + std::to_string(ctx.facts().cu_count) + " hip"
        + std::to_string(rt);
}

std::string tuning_path(const context& ctx, const std::string& name)
{
    if (ctx.cache_dir().empty()) return std::string();
    std::string arch = ctx.facts().arch;
    for (char& c : arch)
        if (c == ':' || c == '/' || c == ' ') c = '_';
    return ctx.cache_dir() + "/clc_" + name + "_" + arch + ".txt";
}

} // namespace

bool load_tuning(const context& ctx, const std::string& name, std::vector<int>& out)
{
    const std::string path = tuning_path(ctx, name);
    return !path.empty() && read_tuning_file(path, tuning_identity(ctx), out);
}

void store_tuning(const context& ctx, const std::string& name,
                  const std::vector<int>& values)
{
    const std::string path = tuning_path(ctx, name);
    if (!path.empty()) (void)write_tuning_file(path, tuning_identity(ctx), values);
}

hipStream_t context::aux_stream(int i)
{
    if (i < 0 || i >= AUX_STREAMS) throw std::out_of_range("clc: auxiliary stream index");
    if (!aux_streams_[i])
        CLC_HIP(hipStreamCreateWithFlags(&aux_streams_[i], hipStreamNonBlocking));
    return aux_streams_[i];
}

hipEvent_t context::aux_event(int i)
{
    if (i < 0 || i >= AUX_EVENTS) throw std::out_of_range("clc: auxiliary event index");
    if (!aux_events_[i])
        CLC_HIP(hipEventCreateWithFlags(&aux_events_[i], hipEventDisableTiming));
    return aux_events_[i];
}

rtc::program_cache& context::programs()
{
    if (!programs_) programs_.reset(new rtc::program_cache(cache_dir_, facts_.arch));
    return *programs_;
}

void context::activate() const
{
    int cur = -1;
    CLC_HIP(hipGetDevice(&cur));
    if (cur != device_) CLC_HIP(hipSetDevice(device_));
}

} // namespace clc
