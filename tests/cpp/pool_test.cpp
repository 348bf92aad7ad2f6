// CPU test of the temporary-storage allocator (opencl-sdk_amd/csrc/pool.cpp)
// against a fake upstream that hands out host addresses and can run out of
// memory on demand. Mirrors the reference's test_memory_resource
// (lib/test/Compute/main.cpp:56-89) and exercises the release-and-retry path
// (MemoryResource.cpp:73-90). Prints "ok N" on success.
#include "../../opencl-sdk_amd/csrc/pool.hpp"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

namespace {

int failures = 0, checks = 0;
#define CHECK(c)                                                               \
    do                                                                         \
    {                                                                          \
        ++checks;                                                              \
        if (!(c))                                                              \
        {                                                                      \
            ++failures;                                                        \
            std::fprintf(stderr, "FAILED: %s at line %d\n", #c, __LINE__);     \
        }                                                                      \
    } while (0)

//! Upstream with a byte budget: allocations beyond it throw out_of_memory.
//! Thread-safe, as every upstream must be: the pool calls it unlocked.
class fake_upstream final : public clc::memory_resource {
public:
    explicit fake_upstream(std::size_t budget) : budget_(budget) {}
    std::size_t live_bytes = 0, calls = 0, frees = 0;
    std::set<void*> live;

protected:
    void* do_allocate(std::size_t bytes) override
    {
        std::lock_guard<std::mutex> lock(m_);
        ++calls;
        if (live_bytes + bytes > budget_) throw clc::out_of_memory("fake: out of memory");
        void* p = std::malloc(bytes);
        live_bytes += bytes;
        live.insert(p);
        return p;
    }
    void do_deallocate(void* p, std::size_t bytes) noexcept override
    {
        std::lock_guard<std::mutex> lock(m_);
        ++frees;
        live_bytes -= bytes;
        live.erase(p);
        std::free(p);
    }

private:
    std::size_t budget_;
    std::mutex m_;
};

void test_size_classes_and_reuse()
{
    auto up = std::make_shared<fake_upstream>(std::size_t(1) << 40);
    auto stats = std::make_shared<clc::statistics_adaptor>(up);
    clc::pool_resource pool(stats);
    CHECK(pool.size_class(1) == 256);
    CHECK(pool.size_class(300) == 512);
    CHECK(pool.size_class((16u << 20) + 1) == (17u << 20));

    void* a = pool.allocate(1000);
    void* b = pool.allocate(1000);
    CHECK(stats->statistics().current_allocations == 2);
    pool.deallocate(a, 1000);
    CHECK(pool.cached_bytes() == 1024);
    void* c = pool.allocate(900); // same class: served from the cache
    CHECK(c == a);
    CHECK(stats->statistics().total_allocations == 2);
    pool.deallocate(b, 1000);
    pool.deallocate(c, 900);
    pool.release();
    CHECK(pool.cached_bytes() == 0);
    CHECK(stats->statistics().current_allocations == 0);
    CHECK(stats->statistics().peak_bytes == 2048);
    CHECK(up->live.empty());
}

void test_out_of_memory_release_and_retry()
{
    // Budget for exactly two 1 MiB blocks.
    auto up = std::make_shared<fake_upstream>(std::size_t(2) << 20);
    clc::pool_resource pool(up);
    void* a = pool.allocate(1 << 20);
    void* b = pool.allocate(1 << 20);
    pool.deallocate(a, 1 << 20);
    pool.deallocate(b, 1 << 20);
    CHECK(pool.cached_bytes() == (std::size_t(2) << 20));
    // A different size class does not fit next to the cached blocks: the
    // pool must release its cache and retry once.
    void* c = pool.allocate(3 << 19);
    CHECK(c != nullptr);
    CHECK(pool.cached_bytes() == 0);
    // Nothing cached and still too big: the error propagates.
    bool threw = false;
    try
    {
        pool.allocate(std::size_t(1) << 21);
    } catch (const clc::out_of_memory&)
    {
        threw = true;
    }
    CHECK(threw);
    pool.deallocate(c, 3 << 19);
}

void test_cache_cap()
{
    auto up = std::make_shared<fake_upstream>(std::size_t(1) << 40);
    clc::pool_resource::options o;
    o.max_cached_bytes = 4096;
    clc::pool_resource pool(up, o);
    void* p[3];
    for (auto& q : p) q = pool.allocate(2048);
    for (auto& q : p) pool.deallocate(q, 2048);
    CHECK(pool.cached_bytes() == 4096); // the third block went back upstream
    CHECK(up->frees == 1);
}

void test_temp_buffer_and_threads()
{
    auto up = std::make_shared<fake_upstream>(std::size_t(1) << 40);
    auto pool = std::make_shared<clc::pool_resource>(up);
    {
        clc::temp_buffer t(pool, 100);
        CHECK(t.get() != nullptr && t.size() == 100);
        clc::temp_buffer z(pool, 0);
        CHECK(z.get() == nullptr);
    }
    CHECK(pool->cached_bytes() == 256);
    std::vector<std::thread> ts;
    for (int t = 0; t < 8; ++t)
        ts.emplace_back([&] {
            for (int i = 0; i < 2000; ++i)
            {
                clc::temp_buffer b(pool, std::size_t(256) << (i % 6));
            }
        });
    for (auto& t : ts) t.join();
    pool->release();
    CHECK(up->live.empty());
}

} // namespace

int main()
{
    test_size_classes_and_reuse();
    test_out_of_memory_release_and_retry();
    test_cache_cap();
    test_temp_buffer_and_threads();
    if (failures)
    {
        std::printf("FAILED %d of %d\n", failures, checks);
        return 1;
    }
    std::printf("ok %d\n", checks);
    return 0;
}
