// The C++ algorithm API (include/compute.hpp) on the GPU, driven the way the
// reference's C++ tests drive theirs (lib/test/Compute/main.cpp: test_reduce
// :182-205, test_scan :207-244, test_sort :246-312,
// test_copy_if_and_histogram :314-337, and the device_buffer part of
// test_memory_resource :56-89): the same scenarios and sizes, each checked
// against a host computation with the standard library. Plus what the
// header adds: a context bound to a caller's hipStream_t, 64-bit sort keys,
// the segmented and learned sorts, and the exceptions that failures raise.
//
// Needs a GPU. Prints "compute_api_test: N checks, 0 failures" on success.
#include "compute.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

namespace cc = cl::compute;

namespace {

int n_checks = 0, n_failed = 0;

void expect(bool ok, const char* what, int line)
{
    ++n_checks;
    if (!ok)
    {
        ++n_failed;
        std::fprintf(stderr, "compute_api_test:%d: %s failed\n", line, what);
    }
}
#define CHECK(c) expect((c), #c, __LINE__)

template <typename E, typename F> bool throws(F&& f)
{
    try
    {
        f();
    } catch (const E&)
    {
        return true;
    } catch (...)
    {
        return false;
    }
    return false;
}

// The reference test's sizes (main.cpp:54).
const std::size_t sizes[] = { 0, 1, 7, 255, 256, 1000, 4097, 100003, std::size_t(1) << 20 };

//! n values uniform in [0, hi] from a splitmix64 stream.
std::vector<uint32_t> uints(std::size_t n, uint32_t hi, uint64_t seed)
{
    std::vector<uint32_t> v(n);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    for (auto& x : v)
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x = hi == 0xFFFFFFFFu ? uint32_t(z) : uint32_t(z % (uint64_t(hi) + 1));
    }
    return v;
}

void test_reduce(cc::compute_context& ctx)
{
    for (std::size_t n : sizes)
    {
        const auto h = uints(n, 1000, n);
        auto b = cc::device_buffer<uint32_t>::from_host(ctx, h);
        CHECK(cc::reduce(ctx, b, cc::ops::plus(), uint32_t(7))
              == std::accumulate(h.begin(), h.end(), uint32_t(7)));
        if (n != 0)
        {
            CHECK(cc::reduce(ctx, b, cc::ops::maximum<uint32_t>(), uint32_t(0))
                  == *std::max_element(h.begin(), h.end()));
            CHECK(cc::reduce(ctx, b, cc::ops::minimum<uint32_t>(), uint32_t(0xFFFFFFFFu))
                  == *std::min_element(h.begin(), h.end()));
        }
    }
    std::vector<float> f(10000);
    for (std::size_t i = 0; i < f.size(); ++i) f[i] = float(i % 17) - 8.0f;
    auto fb = cc::device_buffer<float>::from_host(ctx, f);
    CHECK(cc::reduce(ctx, fb, cc::ops::maximum<float>(), -INFINITY) == 8.0f);
    CHECK(cc::reduce(ctx, fb, cc::ops::minimum<float>(), INFINITY) == -8.0f);
    // A run-time compiled operator (any expression, as binary_op allows).
    CHECK(cc::reduce(ctx, fb, cc::binary_op{ "a + b + 0.0f", "0" }, 0.0f)
          == std::accumulate(f.begin(), f.end(), 0.0f));
}

void test_scan(cc::compute_context& ctx)
{
    for (std::size_t n : sizes)
    {
        const auto h = uints(n, 100, n + 1);
        auto in = cc::device_buffer<uint32_t>::from_host(ctx, h);
        cc::device_buffer<uint32_t> out(ctx, n);
        std::vector<uint32_t> inc(n), exc(n);
        uint32_t run = 0;
        for (std::size_t i = 0; i < n; ++i)
        {
            exc[i] = run;
            inc[i] = run += h[i];
        }
        cc::inclusive_scan(ctx, in, out);
        CHECK(out.to_host(ctx) == inc);
        cc::exclusive_scan(ctx, in, out);
        CHECK(out.to_host(ctx) == exc);
        cc::inclusive_scan(ctx, in, in); // in place
        CHECK(in.to_host(ctx) == inc);
    }
    // The right projection scans to the input itself only if operands are
    // always combined in sequence order (not commutative).
    std::vector<int32_t> v(5000);
    for (std::size_t i = 0; i < v.size(); ++i) v[i] = int32_t(i * 13 % 97);
    auto vb = cc::device_buffer<int32_t>::from_host(ctx, v);
    cc::device_buffer<int32_t> vo(ctx, v.size());
    cc::inclusive_scan(ctx, vb, vo, cc::binary_op{ "b", "0" });
    CHECK(vo.to_host(ctx) == v);
    cc::inclusive_scan(ctx, vb, vo, cc::ops::maximum<int32_t>());
    std::vector<int32_t> mx(v.size());
    int32_t m = INT32_MIN;
    for (std::size_t i = 0; i < v.size(); ++i) mx[i] = m = std::max(m, v[i]);
    CHECK(vo.to_host(ctx) == mx);
}

template <typename K> std::vector<uint32_t> stable_order(const std::vector<K>& keys, bool desc)
{
    std::vector<uint32_t> idx(keys.size());
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        return desc ? keys[b] < keys[a] : keys[a] < keys[b];
    });
    return idx;
}

void test_sort(cc::compute_context& ctx)
{
    for (std::size_t n : sizes)
    {
        const auto keys = uints(n, 0xFFFFFFFFu, n + 2);
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        auto kb = cc::device_buffer<uint32_t>::from_host(ctx, keys);
        auto vb = cc::device_buffer<uint32_t>::from_host(ctx, idx);
        cc::sort_by_key(ctx, kb, vb);
        const auto order = stable_order(keys, false);
        std::vector<uint32_t> rk(n);
        for (std::size_t i = 0; i < n; ++i) rk[i] = keys[order[i]];
        CHECK(kb.to_host(ctx) == rk);
        CHECK(vb.to_host(ctx) == order);
    }
    // Stability with many duplicates.
    const auto dup = uints(50000, 15, 99);
    std::vector<uint32_t> idx(dup.size());
    std::iota(idx.begin(), idx.end(), 0u);
    auto kb = cc::device_buffer<uint32_t>::from_host(ctx, dup);
    auto vb = cc::device_buffer<uint32_t>::from_host(ctx, idx);
    cc::sort_by_key(ctx, kb, vb);
    CHECK(vb.to_host(ctx) == stable_order(dup, false));
    // Signed keys with the extremes.
    std::vector<int32_t> ints{ 5, -3, 0, INT32_MIN, 7, -1, INT32_MAX, 2 };
    auto ib = cc::device_buffer<int32_t>::from_host(ctx, ints);
    cc::sort(ctx, ib);
    auto si = ints;
    std::sort(si.begin(), si.end());
    CHECK(ib.to_host(ctx) == si);
    // Floats descending with -0.0 and both infinities: compared bit for bit,
    // since the sort puts +0.0 before -0.0 when descending.
    std::vector<float> fl(3000);
    const auto fr = uints(fl.size(), 0xFFFFFFFFu, 5);
    for (std::size_t i = 0; i < fl.size(); ++i) fl[i] = (float(fr[i]) / 4294967296.0f - 0.5f) * 2e6f;
    fl[0] = -0.0f;
    fl[1] = INFINITY;
    fl[2] = -INFINITY;
    fl[3] = 0.0f;
    auto fb = cc::device_buffer<float>::from_host(ctx, fl);
    cc::sort(ctx, fb, true);
    std::vector<uint32_t> bits(fl.size());
    for (std::size_t i = 0; i < fl.size(); ++i)
    {
        uint32_t u;
        std::memcpy(&u, &fl[i], 4);
        bits[i] = (u & 0x80000000u) ? ~u : (u | 0x80000000u); // order image
    }
    const auto forder = stable_order(bits, true);
    const auto got = fb.to_host(ctx);
    bool same = got.size() == fl.size();
    for (std::size_t i = 0; same && i < fl.size(); ++i)
        same = std::memcmp(&got[i], &fl[forder[i]], 4) == 0;
    CHECK(same);
    // 64-bit keys (an extension: the reference sorts 32-bit keys only).
    std::vector<uint64_t> wide(70001);
    const auto lo = uints(wide.size(), 0xFFFFFFFFu, 6), hi = uints(wide.size(), 255, 7);
    for (std::size_t i = 0; i < wide.size(); ++i) wide[i] = (uint64_t(hi[i]) << 32) | lo[i];
    std::vector<uint32_t> widx(wide.size());
    std::iota(widx.begin(), widx.end(), 0u);
    auto wb = cc::device_buffer<uint64_t>::from_host(ctx, wide);
    auto wv = cc::device_buffer<uint32_t>::from_host(ctx, widx);
    cc::sort_by_key(ctx, wb, wv, true);
    CHECK(wv.to_host(ctx) == stable_order(wide, true));
}

void test_copy_if_and_histogram(cc::compute_context& ctx)
{
    for (std::size_t n : sizes)
    {
        const auto h = uints(n, 1000, n + 3);
        auto in = cc::device_buffer<uint32_t>::from_host(ctx, h);
        cc::device_buffer<uint32_t> out(ctx, n);
        const std::size_t count = cc::copy_if(ctx, in, out, "(x % 3u) == 0u");
        std::vector<uint32_t> ref;
        for (uint32_t x : h)
            if (x % 3 == 0) ref.push_back(x);
        CHECK(count == ref.size());
        auto got = out.to_host(ctx);
        got.resize(count);
        CHECK(got == ref);
        cc::device_buffer<uint32_t> hist(ctx, 10);
        cc::histogram_even(ctx, in, hist, 0.0, 1000.0);
        std::vector<uint32_t> href(10, 0);
        for (uint32_t x : h)
            if (x < 1000) ++href[x / 100];
        CHECK(hist.to_host(ctx) == href);
    }
}

void test_elementwise(cc::compute_context& ctx)
{
    const std::size_t n = 1000;
    cc::device_buffer<int32_t> seq(ctx, n);
    cc::sequence(ctx, seq, int32_t(-5), int32_t(3));
    std::vector<int32_t> want(n);
    for (std::size_t i = 0; i < n; ++i) want[i] = -5 + 3 * int32_t(i);
    CHECK(seq.to_host(ctx) == want);
    cc::device_buffer<int32_t> prod(ctx, n);
    cc::transform(ctx, seq, seq, prod, "x * y - 1");
    for (auto& w : want) w = w * w - 1;
    CHECK(prod.to_host(ctx) == want);
    cc::device_buffer<float> half(ctx, n);
    cc::transform(ctx, seq, half, "(float)x * 0.5f");
    CHECK(half.to_host(ctx)[3] == 2.0f);
    cc::fill(ctx, half, 2.5f);
    CHECK(half.to_host(ctx) == std::vector<float>(n, 2.5f));
}

void test_buffers_and_errors(cc::compute_context& ctx)
{
    // Move-only; a moved-from buffer is empty; the block (512 B class)
    // returns to the pool once, when its last owner goes.
    clc_memory_stats mid{}, after{};
    {
        cc::device_buffer<float> x(ctx, 100);
        cc::device_buffer<float> y(std::move(x));
        CHECK(x.empty() && x.get() == nullptr && y.size() == 100 && y.get() != nullptr);
        cc::device_buffer<float> z;
        z = std::move(y);
        CHECK(y.empty() && z.size() == 100 && z.bytes() == 400);
        clc_context_memory_stats(ctx.handle(), &mid);
    }
    clc_context_memory_stats(ctx.handle(), &after);
    CHECK(after.cached_bytes == mid.cached_bytes + 512);
    CHECK(after.current_bytes == mid.current_bytes);
    cc::device_buffer<uint32_t> b(ctx, 8);
    std::vector<uint32_t> h(9, 1);
    CHECK(throws<std::out_of_range>([&] { b.copy_from_host(ctx, h.data(), 9); }));
    CHECK(throws<std::out_of_range>([&] { b.copy_to_host(ctx, h.data(), 2, 7); }));
    cc::device_buffer<uint32_t> small(ctx, 4);
    CHECK(throws<std::invalid_argument>([&] { cc::inclusive_scan(ctx, b, small); }));
    CHECK(throws<std::invalid_argument>([&] { cc::sort_by_key(ctx, b, small); }));
    // A compile error in an operator expression carries CLC_ERROR_BUILD.
    bool build_error = false;
    try
    {
        cc::reduce(ctx, b, cc::binary_op{ "a +* b", "0" }, 0u);
    } catch (const cc::error& e)
    {
        build_error = e.status() == CLC_ERROR_BUILD;
    }
    CHECK(build_error);
    // A bitwise operator on floats is refused, as the reference's compiler
    // refuses it.
    cc::device_buffer<float> fb(ctx, 4);
    CHECK(throws<cc::error>([&] { cc::reduce(ctx, fb, cc::ops::bit_or(), 0.0f); }));
}

void test_stream_context_and_extensions()
{
    hipStream_t s = nullptr;
    CHECK(hipStreamCreate(&s) == hipSuccess);
    {
        cc::compute_context ctx(s);
        CHECK(ctx.queue() == s);
        // segmented sort: three segments of 3000, 0 and 70000 keys.
        const auto keys = uints(73000, 0xFFFFFFFFu, 11);
        std::vector<uint32_t> idx(keys.size());
        std::iota(idx.begin(), idx.end(), 0u);
        const std::vector<uint64_t> off{ 0, 3000, 3000, 73000 };
        auto kb = cc::device_buffer<uint32_t>::from_host(ctx, keys);
        auto vb = cc::device_buffer<uint32_t>::from_host(ctx, idx);
        auto ob = cc::device_buffer<uint64_t>::from_host(ctx, off);
        cc::segmented_sort(ctx, kb, &vb, ob);
        std::vector<uint32_t> want(keys.size());
        for (std::size_t sgi = 0; sgi + 1 < off.size(); ++sgi)
        {
            std::vector<uint32_t> part(keys.begin() + off[sgi], keys.begin() + off[sgi + 1]);
            const auto o = stable_order(part, false);
            for (std::size_t i = 0; i < o.size(); ++i) want[off[sgi] + i] = uint32_t(off[sgi]) + o[i];
        }
        CHECK(vb.to_host(ctx) == want);
        // learned sort: same output as sort_by_key.
        std::vector<uint64_t> wk(1 << 20);
        const auto a = uints(wk.size(), 0xFFFFFFFFu, 12), c = uints(wk.size(), 0xFFFFu, 13);
        for (std::size_t i = 0; i < wk.size(); ++i) wk[i] = (uint64_t(c[i]) << 24) ^ a[i];
        std::vector<uint32_t> wi(wk.size());
        std::iota(wi.begin(), wi.end(), 0u);
        auto k1 = cc::device_buffer<uint64_t>::from_host(ctx, wk);
        auto v1 = cc::device_buffer<uint32_t>::from_host(ctx, wi);
        clc_learned_stats st{};
        cc::learned_sort(ctx, k1, &v1, false, &st);
        CHECK(v1.to_host(ctx) == stable_order(wk, false));
        ctx.finish();
    }
    CHECK(hipStreamDestroy(s) == hipSuccess);
}

} // namespace

int main()
{
    {
        cc::compute_context ctx;
        test_reduce(ctx);
        test_scan(ctx);
        test_sort(ctx);
        test_copy_if_and_histogram(ctx);
        test_elementwise(ctx);
        test_buffers_and_errors(ctx);
    }
    test_stream_context_and_extensions();
    std::printf("compute_api_test: %d checks, %d failures\n", n_checks, n_failed);
    return n_failed == 0 ? 0 : 1;
}
