// C++ interface of the MI355X sort / scan / reduce / histogram library: the
// reference's C++ algorithm API (lib/include/CL/Compute/Algorithm.hpp:25-250,
// DeviceBuffer.hpp:26-132, Context.hpp:25-59) as a header-only layer over the
// C ABI in compute_c.h, so C++ callers of the reference (its tests,
// lib/test/Compute/main.cpp:182-337, and the compute-pipeline sample) keep
// their code: the same namespace (cl::compute), names, argument order,
// defaults and stability guarantee.
//
// What changes, all following from HIP replacing OpenCL:
//   - compute_context binds a hipStream_t where the reference binds a
//     cl::CommandQueue (Context.hpp:33-36); it can also create its own
//     stream on a device, with an optional cache directory.
//   - device_buffer<T>::get() is the T* device pointer where the reference
//     returns its cl::Buffer; the block comes from the device's scratch pool
//     (clc_buffer_alloc). copy_from_host / copy_to_host take the context
//     (its stream) where the reference takes a queue.
//   - sort and sort_by_key also accept 64-bit keys (the reference's
//     static_assert allows 32-bit keys only, Algorithm.hpp:206,217).
//   - Failures surface as exceptions, as in the reference: std::invalid_argument
//     for CLC_ERROR_INVALID_ARGUMENT, std::length_error for sizes beyond
//     2^31-1 (check_size, Detail.hpp:18-23), cl::compute::error (with the
//     clc_status and the hipError_t) for everything else.
// Extensions with no reference counterpart: segmented_sort, learned_sort.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "compute_c.h"

namespace cl {
namespace compute {

//! A failed library call: the clc_status, the HIP error (for
//! CLC_ERROR_OPENCL / CLC_ERROR_OUT_OF_MEMORY) and clc_last_error()'s text.
class error : public std::runtime_error {
public:
    error(clc_status status, int hip_error, const std::string& what)
        : std::runtime_error(what), status_(status), hip_error_(hip_error)
    {}
    clc_status status() const { return status_; }
    int hip_error() const { return hip_error_; }

private:
    clc_status status_;
    int hip_error_;
};

namespace detail {

inline void check(clc_status s)
{
    if (s == CLC_SUCCESS) return;
    const std::string what = clc_last_error();
    if (s == CLC_ERROR_INVALID_ARGUMENT)
    {
        if (what.find("2^31-1") != std::string::npos) throw std::length_error(what);
        throw std::invalid_argument(what);
    }
    throw error(s, clc_last_cl_error(), what);
}

//! The clc_type of a host element type (the reference's type_info /
//! type_name, Algorithm.hpp:68-77, ProgramCache.hpp:114-130).
template <typename T> struct type_of;
#define CLC_CXX_TYPE(HOST, CODE)                                               \
    template <> struct type_of<HOST>                                           \
    {                                                                          \
        static constexpr clc_type value = CODE;                                \
    };
CLC_CXX_TYPE(int8_t, CLC_TYPE_I8)
CLC_CXX_TYPE(uint8_t, CLC_TYPE_U8)
CLC_CXX_TYPE(int16_t, CLC_TYPE_I16)
CLC_CXX_TYPE(uint16_t, CLC_TYPE_U16)
CLC_CXX_TYPE(int32_t, CLC_TYPE_I32)
CLC_CXX_TYPE(uint32_t, CLC_TYPE_U32)
CLC_CXX_TYPE(int64_t, CLC_TYPE_I64)
CLC_CXX_TYPE(uint64_t, CLC_TYPE_U64)
CLC_CXX_TYPE(float, CLC_TYPE_F32)
CLC_CXX_TYPE(double, CLC_TYPE_F64)
#undef CLC_CXX_TYPE

} // namespace detail

//! Associative binary operator: an expression of a and b plus its identity
//! (the reference's binary_op, Algorithm.hpp:27-31). The ops:: forms run
//! ahead-of-time kernels; any other expression is compiled once per context.
struct binary_op
{
    std::string expression;
    std::string identity;
};

namespace ops {
//! Lowest and highest values as identity expressions (Algorithm.hpp:33-49).
template <typename T> struct limits;
#define CLC_CXX_LIMITS(HOST, LOWEST, HIGHEST)                                  \
    template <> struct limits<HOST>                                            \
    {                                                                          \
        static const char* lowest() { return LOWEST; }                         \
        static const char* highest() { return HIGHEST; }                       \
    };
CLC_CXX_LIMITS(int8_t, "CHAR_MIN", "CHAR_MAX")
CLC_CXX_LIMITS(uint8_t, "0", "UCHAR_MAX")
CLC_CXX_LIMITS(int16_t, "SHRT_MIN", "SHRT_MAX")
CLC_CXX_LIMITS(uint16_t, "0", "USHRT_MAX")
CLC_CXX_LIMITS(int32_t, "INT_MIN", "INT_MAX")
CLC_CXX_LIMITS(uint32_t, "0", "UINT_MAX")
CLC_CXX_LIMITS(int64_t, "LONG_MIN", "LONG_MAX")
CLC_CXX_LIMITS(uint64_t, "0", "ULONG_MAX")
CLC_CXX_LIMITS(float, "(-INFINITY)", "INFINITY")
CLC_CXX_LIMITS(double, "(-INFINITY)", "INFINITY")
#undef CLC_CXX_LIMITS

// The canonical operators and identities (Algorithm.hpp:51-65): these
// strings are the API.
inline binary_op plus() { return { "(a + b)", "0" }; }
inline binary_op multiplies() { return { "(a * b)", "1" }; }
inline binary_op bit_and() { return { "(a & b)", "(~0)" }; }
inline binary_op bit_or() { return { "(a | b)", "0" }; }
inline binary_op bit_xor() { return { "(a ^ b)", "0" }; }
template <typename T> binary_op minimum()
{
    return { "((b < a) ? b : a)", limits<T>::highest() };
}
template <typename T> binary_op maximum()
{
    return { "((a < b) ? b : a)", limits<T>::lowest() };
}
} // namespace ops

//! One device and one in-order hipStream_t that every algorithm enqueues on,
//! with the device's scratch pool behind it (the reference's compute_context).
//! Not thread-safe: one per thread. Move-only.
class compute_context {
public:
    //! A new stream on the first GPU.
    compute_context() { detail::check(clc_context_create(CLC_DEFAULT_DEVICE, &ctx_)); }
    //! Binds a caller-owned stream on the current device; the stream must
    //! outlive the context.
    explicit compute_context(hipStream_t stream)
    {
        detail::check(clc_context_create_from_stream(static_cast<void*>(stream), &ctx_));
    }
    //! A new stream on device index; cache_dir keeps run-time compiled
    //! operators and tuned launch tables across processes ("" = the
    //! CL_COMPUTE_CACHE_DIR environment variable, if set).
    compute_context(std::size_t index, const std::string& cache_dir)
    {
        detail::check(clc_context_create_with_cache(index, cache_dir.c_str(), &ctx_));
    }
    compute_context(const compute_context&) = delete;
    compute_context& operator=(const compute_context&) = delete;
    compute_context(compute_context&& o) noexcept : ctx_(o.ctx_) { o.ctx_ = nullptr; }
    compute_context& operator=(compute_context&& o) noexcept
    {
        if (this != &o)
        {
            clc_context_release(ctx_);
            ctx_ = o.ctx_;
            o.ctx_ = nullptr;
        }
        return *this;
    }
    ~compute_context() { clc_context_release(ctx_); }

    clc_context handle() const { return ctx_; }
    //! The stream the algorithms enqueue on (the reference's queue()).
    hipStream_t queue() const
    {
        void* s = nullptr;
        detail::check(clc_context_get_stream(ctx_, &s));
        return static_cast<hipStream_t>(s);
    }
    //! Waits for everything queued on the context's stream.
    void finish() const { detail::check(clc_finish(ctx_)); }

private:
    clc_context ctx_ = nullptr;
};

//! Typed, move-only device allocation from the device's scratch pool (the
//! reference's device_buffer, DeviceBuffer.hpp:26-132). The block may be
//! larger than size() * sizeof(T); algorithms are always given size().
template <typename T> class device_buffer {
public:
    device_buffer() = default;
    device_buffer(const compute_context& ctx, std::size_t size) : size_(size)
    {
        if (size_ != 0) detail::check(clc_buffer_alloc(ctx.handle(), bytes(), &buf_));
    }
    device_buffer(const device_buffer&) = delete;
    device_buffer& operator=(const device_buffer&) = delete;
    device_buffer(device_buffer&& o) noexcept : buf_(o.buf_), size_(o.size_)
    {
        o.buf_ = nullptr;
        o.size_ = 0;
    }
    device_buffer& operator=(device_buffer&& o) noexcept
    {
        if (this != &o)
        {
            reset();
            buf_ = o.buf_;
            size_ = o.size_;
            o.buf_ = nullptr;
            o.size_ = 0;
        }
        return *this;
    }
    ~device_buffer() { reset(); }

    //! Returns the block to the pool; work already queued on the context's
    //! stream may still use it (the pool hands it out again only to work
    //! ordered after).
    void reset() noexcept
    {
        clc_buffer_release(buf_);
        buf_ = nullptr;
        size_ = 0;
    }

    std::size_t size() const { return size_; }
    std::size_t bytes() const { return size_ * sizeof(T); }
    bool empty() const { return size_ == 0; }
    //! The device pointer (nullptr when empty).
    T* get() const { return static_cast<T*>(clc_buffer_data(buf_)); }
    clc_buffer handle() const { return buf_; }

    //! Copies count elements from host src to [offset, offset + count); the
    //! range must lie within the buffer (std::out_of_range otherwise).
    //! Blocking.
    void copy_from_host(compute_context& ctx, const T* src, std::size_t count,
                        std::size_t offset = 0)
    {
        check_range(offset, count);
        if (count != 0)
            detail::check(clc_buffer_write(ctx.handle(), buf_, offset * sizeof(T),
                                           count * sizeof(T), src));
    }
    //! Copies [offset, offset + count) to host dst. Blocking.
    void copy_to_host(compute_context& ctx, T* dst, std::size_t count,
                      std::size_t offset = 0) const
    {
        check_range(offset, count);
        if (count != 0)
            detail::check(clc_buffer_read(ctx.handle(), buf_, offset * sizeof(T),
                                          count * sizeof(T), dst));
    }
    std::vector<T> to_host(compute_context& ctx) const
    {
        std::vector<T> out(size_);
        copy_to_host(ctx, out.data(), size_);
        return out;
    }
    static device_buffer from_host(compute_context& ctx, const std::vector<T>& values)
    {
        device_buffer b(ctx, values.size());
        b.copy_from_host(ctx, values.data(), values.size());
        return b;
    }

private:
    void check_range(std::size_t offset, std::size_t count) const
    {
        if (offset > size_ || count > size_ - offset)
            throw std::out_of_range("device_buffer: range out of bounds");
    }

    clc_buffer buf_ = nullptr;
    std::size_t size_ = 0;
};

// ---- algorithms (Algorithm.hpp:119-250) --------------------------------------

//! out[i] = value.
template <typename T> void fill(compute_context& ctx, device_buffer<T>& out, T value)
{
    detail::check(clc_fill(ctx.handle(), out.get(), out.size(), detail::type_of<T>::value,
                           &value));
}

//! out[i] = start + i * step.
template <typename T>
void sequence(compute_context& ctx, device_buffer<T>& out, T start = T(0), T step = T(1))
{
    detail::check(clc_sequence(ctx.handle(), out.get(), out.size(),
                               detail::type_of<T>::value, &start, &step));
}

//! out[i] = expression(x = in[i]).
template <typename In, typename Out>
void transform(compute_context& ctx, const device_buffer<In>& in, device_buffer<Out>& out,
               const std::string& expression)
{
    if (out.size() < in.size()) throw std::invalid_argument("transform: output too small");
    detail::check(clc_transform(ctx.handle(), in.get(), detail::type_of<In>::value, out.get(),
                                detail::type_of<Out>::value, in.size(), expression.c_str()));
}

//! out[i] = expression(x = in1[i], y = in2[i]); out may alias an input.
template <typename In1, typename In2, typename Out>
void transform(compute_context& ctx, const device_buffer<In1>& in1,
               const device_buffer<In2>& in2, device_buffer<Out>& out,
               const std::string& expression)
{
    if (in2.size() < in1.size() || out.size() < in1.size())
        throw std::invalid_argument("transform: size mismatch");
    detail::check(clc_transform2(ctx.handle(), in1.get(), detail::type_of<In1>::value,
                                 in2.get(), detail::type_of<In2>::value, out.get(),
                                 detail::type_of<Out>::value, in1.size(),
                                 expression.c_str()));
}

//! op(init, fold of in); the order of combination is unspecified, so op must
//! be associative and commutative. Blocks until the result is on the host.
template <typename T>
T reduce(compute_context& ctx, const device_buffer<T>& in, const binary_op& op = ops::plus(),
         T init = T(0))
{
    T result;
    detail::check(clc_reduce(ctx.handle(), in.get(), in.size(), detail::type_of<T>::value,
                             op.expression.c_str(), op.identity.c_str(), &init, &result));
    return result;
}

//! out[i] = in[0] op ... op in[i]; out may alias in.
template <typename T>
void inclusive_scan(compute_context& ctx, const device_buffer<T>& in, device_buffer<T>& out,
                    const binary_op& op = ops::plus())
{
    if (out.size() < in.size())
        throw std::invalid_argument("inclusive_scan: output too small");
    detail::check(clc_scan(ctx.handle(), in.get(), out.get(), in.size(),
                           detail::type_of<T>::value, op.expression.c_str(),
                           op.identity.c_str(), 0));
}

//! out[i] = identity op in[0] op ... op in[i-1]; out may alias in.
template <typename T>
void exclusive_scan(compute_context& ctx, const device_buffer<T>& in, device_buffer<T>& out,
                    const binary_op& op = ops::plus())
{
    if (out.size() < in.size())
        throw std::invalid_argument("exclusive_scan: output too small");
    detail::check(clc_scan(ctx.handle(), in.get(), out.get(), in.size(),
                           detail::type_of<T>::value, op.expression.c_str(),
                           op.identity.c_str(), 1));
}

//! Stable in-place radix sort of 32- or 64-bit keys (signed, unsigned or
//! floating point; -0.0 before +0.0).
template <typename Key>
void sort(compute_context& ctx, device_buffer<Key>& keys, bool descending = false)
{
    static_assert(sizeof(Key) == 4 || sizeof(Key) == 8, "sort: keys must be 32- or 64-bit");
    detail::check(clc_sort(ctx.handle(), keys.get(), detail::type_of<Key>::value, nullptr,
                           keys.size(), descending ? 1 : 0));
}

//! Sorts keys and permutes 32-bit values alongside, stably.
template <typename Key, typename Value>
void sort_by_key(compute_context& ctx, device_buffer<Key>& keys, device_buffer<Value>& values,
                 bool descending = false)
{
    static_assert(sizeof(Key) == 4 || sizeof(Key) == 8,
                  "sort_by_key: keys must be 32- or 64-bit");
    static_assert(sizeof(Value) == 4, "sort_by_key: values must be 32-bit");
    if (values.size() < keys.size()) throw std::invalid_argument("sort_by_key: too few values");
    detail::check(clc_sort(ctx.handle(), keys.get(), detail::type_of<Key>::value,
                           values.get(), keys.size(), descending ? 1 : 0));
}

//! Copies the elements satisfying predicate (an expression of x) to the
//! front of out in order; returns how many (blocks until the count is known).
template <typename T>
std::size_t copy_if(compute_context& ctx, const device_buffer<T>& in, device_buffer<T>& out,
                    const std::string& predicate)
{
    if (out.size() < in.size()) throw std::invalid_argument("copy_if: output too small");
    std::size_t count = 0;
    detail::check(clc_copy_if(ctx.handle(), in.get(), out.get(), in.size(),
                              detail::type_of<T>::value, predicate.c_str(), &count));
    return count;
}

//! Counts samples into histogram.size() equal-width bins over [lower, upper);
//! samples outside are ignored.
template <typename T>
void histogram_even(compute_context& ctx, const device_buffer<T>& in,
                    device_buffer<uint32_t>& histogram, double lower, double upper)
{
    detail::check(clc_histogram_even(ctx.handle(), in.get(), in.size(),
                                     detail::type_of<T>::value, histogram.get(),
                                     histogram.size(), lower, upper));
}

// ---- extensions (no reference counterpart) -----------------------------------

//! Stable sort of every segment [offsets[i], offsets[i+1]) of keys (and of
//! values, when given), independently; offsets has one more entry than
//! there are segments.
template <typename Key>
void segmented_sort(compute_context& ctx, device_buffer<Key>& keys,
                    device_buffer<uint32_t>* values, const device_buffer<uint64_t>& offsets,
                    bool descending = false)
{
    if (values != nullptr && values->size() < keys.size())
        throw std::invalid_argument("segmented_sort: too few values");
    const std::size_t nseg = offsets.size() == 0 ? 0 : offsets.size() - 1;
    detail::check(clc_segmented_sort(ctx.handle(), keys.get(), detail::type_of<Key>::value,
                                     values != nullptr ? values->get() : nullptr, keys.size(),
                                     offsets.get(), nseg, descending ? 1 : 0));
}

//! Learned-bucket sort: the same output as sort_by_key (or sort), faster on
//! skewed keys; stats (optional) receives the model's figures.
template <typename Key>
void learned_sort(compute_context& ctx, device_buffer<Key>& keys,
                  device_buffer<uint32_t>* values, bool descending = false,
                  clc_learned_stats* stats = nullptr)
{
    if (values != nullptr && values->size() < keys.size())
        throw std::invalid_argument("learned_sort: too few values");
    detail::check(clc_learned_sort(ctx.handle(), keys.get(), detail::type_of<Key>::value,
                                   values != nullptr ? values->get() : nullptr, keys.size(),
                                   descending ? 1 : 0, stats));
}

} // namespace compute
} // namespace cl
