// Pool and statistics resources; behaviour follows the reference's
// lib/src/Compute/MemoryResource.cpp:27-167 (size classes :47-58, cache hit
// :60-72, release-and-retry on out-of-memory :73-90, capped caching
// :92-111, counters :139-162).
#include "pool.hpp"

#include <algorithm>

namespace clc {

pool_resource::pool_resource(std::shared_ptr<memory_resource> upstream)
    : pool_resource(std::move(upstream), options())
{}

pool_resource::pool_resource(std::shared_ptr<memory_resource> upstream,
                             options opts)
    : upstream_(std::move(upstream)), opts_(opts)
{
    if (!upstream_) throw std::invalid_argument("pool_resource: null upstream");
    if (opts_.min_block_bytes == 0 || opts_.large_granularity == 0)
        throw std::invalid_argument("pool_resource: zero block size");
}

pool_resource::~pool_resource()
{
    std::lock_guard<std::mutex> lock(mutex_);
    release_locked();
}

std::size_t pool_resource::size_class(std::size_t bytes) const
{
    bytes = std::max(bytes, opts_.min_block_bytes);
    if (bytes <= opts_.pow2_limit)
    {
        std::size_t size = 1;
        while (size < bytes) size <<= 1;
        return size;
    }
    const std::size_t g = opts_.large_granularity;
    return (bytes + g - 1) / g * g;
}

void* pool_resource::do_allocate(std::size_t bytes)
{
    const std::size_t size = size_class(bytes);
    {
        std::lock_guard<std::mutex> lock(mutex_);
        auto it = free_lists_.find(size);
        if (it != free_lists_.end() && !it->second.empty())
        {
            void* p = it->second.back();
            it->second.pop_back();
            cached_bytes_ -= size;
            return p;
        }
    }
    try
    {
        return upstream_->allocate(size);
    } catch (const out_of_memory&)
    {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            if (cached_bytes_ == 0) throw;
            release_locked();
        }
        return upstream_->allocate(size);
    }
}

void pool_resource::do_deallocate(void* p, std::size_t bytes) noexcept
{
    const std::size_t size = size_class(bytes);
    std::lock_guard<std::mutex> lock(mutex_);
    if (cached_bytes_ + size <= opts_.max_cached_bytes)
    {
        try
        {
            free_lists_[size].push_back(p);
            cached_bytes_ += size;
            return;
        } catch (...)
        {
            // Out of host memory: hand the block back instead.
        }
    }
    upstream_->deallocate(p, size);
}

void pool_resource::release()
{
    std::lock_guard<std::mutex> lock(mutex_);
    release_locked();
}

void pool_resource::release_locked()
{
    for (auto& entry : free_lists_)
        for (void* p : entry.second) upstream_->deallocate(p, entry.first);
    free_lists_.clear();
    cached_bytes_ = 0;
}

std::size_t pool_resource::cached_bytes() const
{
    std::lock_guard<std::mutex> lock(mutex_);
    return cached_bytes_;
}

statistics_adaptor::statistics_adaptor(
    std::shared_ptr<memory_resource> upstream)
    : upstream_(std::move(upstream))
{
    if (!upstream_)
        throw std::invalid_argument("statistics_adaptor: null upstream");
}

void* statistics_adaptor::do_allocate(std::size_t bytes)
{
    void* p = upstream_->allocate(bytes);
    std::lock_guard<std::mutex> lock(mutex_);
    counters_.current_bytes += bytes;
    counters_.peak_bytes = std::max(counters_.peak_bytes, counters_.current_bytes);
    counters_.total_bytes += bytes;
    ++counters_.current_allocations;
    ++counters_.total_allocations;
    return p;
}

void statistics_adaptor::do_deallocate(void* p, std::size_t bytes) noexcept
{
    upstream_->deallocate(p, bytes);
    std::lock_guard<std::mutex> lock(mutex_);
    counters_.current_bytes -= bytes;
    --counters_.current_allocations;
}

statistics_adaptor::counters statistics_adaptor::statistics() const
{
    std::lock_guard<std::mutex> lock(mutex_);
    return counters_;
}

} // namespace clc
