// Scratch memory for the library's launches: one cache of device blocks per
// GPU, shared by every context on that GPU, that remembers which stream last
// used each block.
//
// Plays the role of the reference's temporary-storage resources
// (lib/include/CL/Compute/MemoryResource.hpp:31-150, Detail.hpp:37-57), which
// recycle a freed block at once because OpenCL queue order makes that safe
// for one in-order queue. Here work runs on several HIP streams: each
// context's own stream, the auxiliary streams an operation forks onto, and
// the streams of other contexts (one per thread) on the same GPU. So a block
// is parked with its owner stream, the stream its last user queued work on:
//   - taken again on the owner stream: reused at once (the stream is in order);
//   - taken on another stream: if the owner stream is idle, at once; else the
//     taker stream is made to wait (on the device, no host wait) for
//     everything queued on the owner stream so far, which includes every use
//     of the block, since those were queued before the block was parked;
//   - parked with no owner (any_stream): the caller promises that no queued
//     work touches it any more.
// A stream that a context is about to destroy is retired first: the pool
// waits for it and drops it as an owner.
//
// Sizes round up to classes: powers of two from 256 B to 16 MiB, then whole
// MiB. The cache keeps parked blocks up to a byte limit and evicts the least
// recently parked ones first. A request the device cannot meet empties the
// cache and is tried once more. The pool counts what it holds from the device
// (bytes and blocks: current, peak, total) and how it served requests.
//
// The device is reached through device_link, so this file has no HIP
// dependency and the policy is testable on a CPU with a fake link.
#pragma once

#include <cstddef>
#include <cstdint>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

namespace clc {

//! Thrown by a device_link that has no memory left.
struct out_of_memory : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

//! An opaque stream handle: a hipStream_t for the device link. A null
//! hipStream_t is a real stream (the legacy default stream), so "no stream"
//! has its own sentinel, any_stream.
using stream_ref = void*;
inline stream_ref const any_stream = reinterpret_cast<stream_ref>(~std::uintptr_t(0));

//! What the pool needs from a device.
class device_link {
public:
    virtual ~device_link() = default;
    //! New device memory; throws out_of_memory when there is none.
    virtual void* obtain(std::size_t bytes) = 0;
    //! Returns memory that no queued work uses any more.
    virtual void surrender(void* p, std::size_t bytes) noexcept = 0;
    //! True when everything queued on s has finished.
    virtual bool idle(stream_ref s) = 0;
    //! Work queued on later from now on waits for the work queued on
    //! earlier so far (a device-side dependency).
    virtual void follow(stream_ref later, stream_ref earlier) = 0;
    //! The host waits until everything queued on s has finished.
    virtual void drain(stream_ref s) noexcept = 0;
};

//! Where launches draw scratch from: the per-GPU pool, or a fixed block that
//! a captured graph owns.
class scratch_source {
public:
    virtual ~scratch_source() = default;
    //! A block of at least bytes that work queued on s from now on may use.
    virtual void* take(std::size_t bytes, stream_ref s) = 0;
    //! Hands p back; owner is the stream on which its last use was queued.
    virtual void give(void* p, std::size_t bytes, stream_ref owner) noexcept = 0;
};

class stream_pool final : public scratch_source {
public:
    struct limits
    {
        std::size_t smallest = 256;                 //!< smallest class
        std::size_t pow2_until = std::size_t(16) << 20; //!< last power-of-two class
        std::size_t granule = std::size_t(1) << 20; //!< step of the larger classes
        std::size_t cache_cap = std::size_t(8) << 30; //!< parked bytes kept at most
    };
    struct figures
    {
        std::size_t held_bytes = 0;     //!< held from the device now (in use + parked)
        std::size_t held_peak = 0;      //!< largest held_bytes so far
        std::size_t obtained_bytes = 0; //!< every byte ever obtained from the device
        std::size_t held_blocks = 0;    //!< blocks held from the device now
        std::size_t obtained_blocks = 0; //!< blocks ever obtained from the device
        std::size_t parked_bytes = 0;   //!< parked (cached) now
        std::size_t reused = 0;         //!< takes served from the cache
        std::size_t cross_stream = 0;   //!< of those, blocks parked by another stream
        std::size_t waits = 0;          //!< of those, device-side waits inserted
    };

    explicit stream_pool(std::shared_ptr<device_link> link);
    stream_pool(std::shared_ptr<device_link> link, limits l);
    ~stream_pool() override;
    stream_pool(const stream_pool&) = delete;
    stream_pool& operator=(const stream_pool&) = delete;

    void* take(std::size_t bytes, stream_ref s) override;
    void give(void* p, std::size_t bytes, stream_ref owner) noexcept override;

    //! The size class a request of bytes is served from.
    std::size_t class_of(std::size_t bytes) const;
    //! Returns every parked block to the device (waiting for owner streams).
    void trim();
    //! Waits for s and forgets it as an owner: its parked blocks become free
    //! for any stream. Call before destroying a stream the pool has seen.
    void retire(stream_ref s) noexcept;
    figures tally() const;
    std::size_t parked_bytes() const;
    const limits& bounds() const { return limits_; }

private:
    struct parked
    {
        void* p;
        std::size_t size;
        stream_ref owner;
    };
    // Parked blocks in the order they were parked (oldest first); each shelf
    // indexes its blocks in that order, so eviction pops the global front and
    // a take scans one shelf from its newest entry.
    using order_list = std::list<parked>;
    using shelf = std::list<order_list::iterator>;

    void* take_parked(std::size_t size, stream_ref s);
    void unpark(shelf& sh, shelf::iterator it);
    void evict_until(std::size_t room);
    void surrender_all();

    std::shared_ptr<device_link> link_;
    limits limits_;
    mutable std::mutex m_;
    order_list order_;
    std::unordered_map<std::size_t, shelf> shelves_;
    figures fig_;
};

//! RAII scratch: taken on construction and handed back on destruction with
//! the stream it was taken for as the owner (every use of it must be queued
//! on that stream, or joined into it, before the destructor runs). Zero bytes
//! takes nothing.
class temp_buffer {
public:
    temp_buffer(std::shared_ptr<scratch_source> src, std::size_t bytes, stream_ref s)
        : src_(std::move(src)), bytes_(bytes), stream_(s)
    {
        if (bytes_ != 0) ptr_ = src_->take(bytes_, stream_);
    }
    ~temp_buffer()
    {
        if (ptr_ != nullptr) src_->give(ptr_, bytes_, stream_);
    }
    temp_buffer(const temp_buffer&) = delete;
    temp_buffer& operator=(const temp_buffer&) = delete;

    void* get() const { return ptr_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    std::size_t size() const { return bytes_; }

private:
    std::shared_ptr<scratch_source> src_;
    std::size_t bytes_;
    stream_ref stream_;
    void* ptr_ = nullptr;
};

} // namespace clc
