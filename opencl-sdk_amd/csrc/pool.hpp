// Temporary-storage allocator: stream-ordered device memory resources.
//
// Mirrors the reference's memory_resource family
// (lib/include/CL/Compute/MemoryResource.hpp:31-150,
// lib/src/Compute/MemoryResource.cpp:12-167) over raw device pointers:
//   memory_resource     abstract allocate/deallocate          (:31-50)
//   pool_resource       size classes, free lists, cache cap,
//                       release-and-retry on out-of-memory    (:63-120)
//   statistics_adaptor  current/peak/total counters           (:122-150)
// The upstream that talks to HIP lives in device_resource.hpp; this header
// has no HIP dependency so the pool logic is testable on a CPU with a fake
// upstream.
//
// Allocations are stream-ordered, as the reference's are queue-ordered
// (MemoryResource.hpp:25-29): a block handed back by deallocate() may be
// returned by the next allocate() at once, which is safe because every
// command touching it was enqueued on the same in-order stream.
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace clc {

//! Thrown by an upstream resource that has no memory left; the pool
//! releases its cache and retries once when it sees this.
struct out_of_memory : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

class memory_resource {
public:
    virtual ~memory_resource() = default;
    void* allocate(std::size_t bytes) { return do_allocate(bytes); }
    void deallocate(void* p, std::size_t bytes) noexcept
    {
        do_deallocate(p, bytes);
    }

protected:
    virtual void* do_allocate(std::size_t bytes) = 0;
    virtual void do_deallocate(void* p, std::size_t bytes) noexcept = 0;
};

//! Caching pool over an upstream resource (pool_resource,
//! MemoryResource.hpp:63-120). Requests round up to a size class: powers of
//! two from min_block_bytes up to pow2_limit, then multiples of
//! large_granularity. Freed blocks stay in per-class free lists until
//! max_cached_bytes is reached. Thread-safe.
class pool_resource final : public memory_resource {
public:
    struct options
    {
        std::size_t min_block_bytes = 256;
        std::size_t pow2_limit = std::size_t(16) << 20;
        std::size_t large_granularity = std::size_t(1) << 20;
        // The reference caches up to 1 GiB (MemoryResource.hpp:85); a 2^28
        // key-value sort here needs 2 GiB of alternate buffers, and an
        // MI355X has 288 GB, so the default cap is 8 GiB (contexts raise
        // it to an eighth of device memory).
        std::size_t max_cached_bytes = std::size_t(8) << 30;
    };

    explicit pool_resource(std::shared_ptr<memory_resource> upstream);
    pool_resource(std::shared_ptr<memory_resource> upstream, options opts);
    ~pool_resource() override;

    //! Returns every cached block to the upstream resource.
    void release();
    std::size_t cached_bytes() const;
    std::size_t size_class(std::size_t bytes) const;

protected:
    void* do_allocate(std::size_t bytes) override;
    void do_deallocate(void* p, std::size_t bytes) noexcept override;

private:
    void release_locked();

    std::shared_ptr<memory_resource> upstream_;
    options opts_;
    mutable std::mutex mutex_;
    std::map<std::size_t, std::vector<void*>> free_lists_;
    std::size_t cached_bytes_ = 0;
};

//! Counts what passes through it (statistics_adaptor,
//! MemoryResource.hpp:122-150). Thread-safe.
class statistics_adaptor final : public memory_resource {
public:
    struct counters
    {
        std::size_t current_bytes = 0;
        std::size_t peak_bytes = 0;
        std::size_t total_bytes = 0;
        std::size_t current_allocations = 0;
        std::size_t total_allocations = 0;
    };

    explicit statistics_adaptor(std::shared_ptr<memory_resource> upstream);
    counters statistics() const;

protected:
    void* do_allocate(std::size_t bytes) override;
    void do_deallocate(void* p, std::size_t bytes) noexcept override;

private:
    std::shared_ptr<memory_resource> upstream_;
    mutable std::mutex mutex_;
    counters counters_;
};

//! RAII scratch drawn from a resource (internal::temp_buffer,
//! lib/src/Compute/Detail.hpp:37-57). Zero bytes allocates nothing.
class temp_buffer {
public:
    temp_buffer(std::shared_ptr<memory_resource> resource, std::size_t bytes)
        : resource_(std::move(resource)), bytes_(bytes)
    {
        if (bytes_ != 0) ptr_ = resource_->allocate(bytes_);
    }
    ~temp_buffer()
    {
        if (bytes_ != 0) resource_->deallocate(ptr_, bytes_);
    }
    temp_buffer(const temp_buffer&) = delete;
    temp_buffer& operator=(const temp_buffer&) = delete;

    void* get() const { return ptr_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    std::size_t size() const { return bytes_; }

private:
    std::shared_ptr<memory_resource> resource_;
    void* ptr_ = nullptr;
    std::size_t bytes_;
};

} // namespace clc
