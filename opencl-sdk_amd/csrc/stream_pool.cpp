// Per-GPU stream-ordered scratch cache (stream_pool.hpp).
#include "stream_pool.hpp"

#include <algorithm>
#include <iterator>

namespace clc {

stream_pool::stream_pool(std::shared_ptr<device_link> link)
    : stream_pool(std::move(link), limits())
{}

stream_pool::stream_pool(std::shared_ptr<device_link> link, limits l)
    : link_(std::move(link)), limits_(l)
{
    if (!link_) throw std::invalid_argument("stream_pool needs a device link");
    if (limits_.smallest == 0 || limits_.granule == 0
        || limits_.pow2_until < limits_.smallest)
        throw std::invalid_argument("stream_pool: size classes need a nonzero smallest "
                                    "class and granule, and pow2_until >= smallest");
}

stream_pool::~stream_pool()
{
    std::lock_guard<std::mutex> lock(m_);
    surrender_all();
}

std::size_t stream_pool::class_of(std::size_t bytes) const
{
    const std::size_t want = std::max(bytes, limits_.smallest);
    if (want <= limits_.pow2_until)
    {
        // Next power of two at or above want (want >= 1 here).
        const unsigned lead = static_cast<unsigned>(__builtin_clzll(
            static_cast<unsigned long long>(want - 1) | 1ull));
        return want == 1 ? 1 : std::size_t(1) << (64 - lead);
    }
    const std::size_t g = limits_.granule;
    const std::size_t whole = want / g + (want % g != 0);
    if (whole > ~std::size_t(0) / g) throw out_of_memory("stream_pool: request too large");
    return whole * g;
}

void* stream_pool::take(std::size_t bytes, stream_ref s)
{
    const std::size_t size = class_of(bytes);
    {
        std::lock_guard<std::mutex> lock(m_);
        if (void* p = take_parked(size, s)) return p;
    }
    void* p = nullptr;
    try
    {
        p = link_->obtain(size);
    } catch (const out_of_memory&)
    {
        // Parked blocks may be what fills the device: return them all and
        // ask once more; with nothing parked the failure stands.
        {
            std::lock_guard<std::mutex> lock(m_);
            if (order_.empty()) throw;
            surrender_all();
        }
        p = link_->obtain(size);
    }
    std::lock_guard<std::mutex> lock(m_);
    fig_.held_bytes += size;
    fig_.held_peak = std::max(fig_.held_peak, fig_.held_bytes);
    fig_.obtained_bytes += size;
    ++fig_.held_blocks;
    ++fig_.obtained_blocks;
    return p;
}

void* stream_pool::take_parked(std::size_t size, stream_ref s)
{
    auto found = shelves_.find(size);
    if (found == shelves_.end() || found->second.empty()) return nullptr;
    shelf& sh = found->second;
    // Newest first: a block this stream parked itself (or one free for
    // anyone) needs no ordering, and the newest is the likeliest to still
    // sit in L2 and the TLB.
    for (auto r = sh.rbegin(); r != sh.rend(); ++r)
    {
        const stream_ref owner = (*r)->owner;
        if (owner == s || owner == any_stream)
        {
            void* p = (*r)->p;
            unpark(sh, std::next(r).base());
            ++fig_.reused;
            return p;
        }
    }
    // Every block here was parked by another stream: take the oldest, the
    // likeliest to be finished, and order this stream after its owner.
    const shelf::iterator oldest = sh.begin();
    const stream_ref owner = (*oldest)->owner;
    if (!link_->idle(owner))
    {
        if (s == any_stream)
            link_->drain(owner); // the taker queues on no stream: wait here
        else
            link_->follow(s, owner);
        ++fig_.waits;
    }
    void* p = (*oldest)->p;
    unpark(sh, oldest);
    ++fig_.reused;
    ++fig_.cross_stream;
    return p;
}

void stream_pool::unpark(shelf& sh, shelf::iterator it)
{
    fig_.parked_bytes -= (*it)->size;
    order_.erase(*it);
    sh.erase(it);
}

void stream_pool::give(void* p, std::size_t bytes, stream_ref owner) noexcept
{
    if (p == nullptr) return;
    std::size_t size;
    try
    {
        size = class_of(bytes);
    } catch (...)
    {
        size = bytes;
    }
    std::lock_guard<std::mutex> lock(m_);
    if (size <= limits_.cache_cap)
    {
        try
        {
            evict_until(size);
            order_.push_back(parked{ p, size, owner });
            try
            {
                shelves_[size].push_back(std::prev(order_.end()));
            } catch (...)
            {
                order_.pop_back();
                throw;
            }
            fig_.parked_bytes += size;
            return;
        } catch (...)
        {
            // No host memory for the bookkeeping: return the block instead.
        }
    }
    if (owner != any_stream) link_->drain(owner);
    link_->surrender(p, size);
    fig_.held_bytes -= size;
    --fig_.held_blocks;
}

void stream_pool::evict_until(std::size_t room)
{
    while (!order_.empty() && fig_.parked_bytes + room > limits_.cache_cap)
    {
        const parked victim = order_.front();
        shelf& sh = shelves_[victim.size];
        // The oldest parked block is the first entry of its own shelf.
        unpark(sh, sh.begin());
        if (victim.owner != any_stream) link_->drain(victim.owner);
        link_->surrender(victim.p, victim.size);
        fig_.held_bytes -= victim.size;
        --fig_.held_blocks;
    }
}

void stream_pool::trim()
{
    std::lock_guard<std::mutex> lock(m_);
    surrender_all();
}

void stream_pool::surrender_all()
{
    // Waiting on an owner that is already idle costs one query, so owners
    // are drained per block rather than collected first.
    for (const parked& b : order_)
    {
        if (b.owner != any_stream) link_->drain(b.owner);
        link_->surrender(b.p, b.size);
        fig_.held_bytes -= b.size;
        --fig_.held_blocks;
    }
    order_.clear();
    shelves_.clear();
    fig_.parked_bytes = 0;
}

void stream_pool::retire(stream_ref s) noexcept
{
    if (s == any_stream) return;
    std::lock_guard<std::mutex> lock(m_);
    bool waited = false;
    for (parked& b : order_)
        if (b.owner == s)
        {
            if (!waited) link_->drain(s);
            waited = true;
            b.owner = any_stream;
        }
}

stream_pool::figures stream_pool::tally() const
{
    std::lock_guard<std::mutex> lock(m_);
    return fig_;
}

std::size_t stream_pool::parked_bytes() const
{
    std::lock_guard<std::mutex> lock(m_);
    return fig_.parked_bytes;
}

} // namespace clc
