// CPU test of the per-GPU scratch pool (opencl-sdk_amd/csrc/stream_pool.cpp)
// against a fake device link: memory is host memory under a byte budget
// (requests beyond it throw out_of_memory), and streams are small integers
// whose "busy" flag the test sets, so the pool's ordering decisions (reuse at
// once, a device-side wait, a host wait) can be observed. Covers the
// behaviour the reference's allocator test pins (size classes, reuse of a
// freed block, counters, release; lib/test/Compute/main.cpp:56-89) and the
// release-and-retry of MemoryResource.cpp's pool on out-of-memory, plus what
// is new here: owner streams, eviction order, retirement and threads.
// Prints "ok N" on success.
#include "../../opencl-sdk_amd/csrc/stream_pool.hpp"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <set>
#include <thread>
#include <utility>
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

clc::stream_ref stream(std::uintptr_t i) { return reinterpret_cast<clc::stream_ref>(i); }

class fake_link final : public clc::device_link {
public:
    explicit fake_link(std::size_t budget) : budget_(budget) {}

    void* obtain(std::size_t bytes) override
    {
        std::lock_guard<std::mutex> lock(m_);
        ++obtains;
        if (live_bytes + bytes > budget_) throw clc::out_of_memory("fake: out of memory");
        void* p = std::malloc(bytes);
        live_bytes += bytes;
        live.insert(p);
        return p;
    }
    void surrender(void* p, std::size_t bytes) noexcept override
    {
        std::lock_guard<std::mutex> lock(m_);
        ++surrenders;
        live_bytes -= bytes;
        live.erase(p);
        std::free(p);
    }
    bool idle(clc::stream_ref s) override
    {
        std::lock_guard<std::mutex> lock(m_);
        return busy.count(s) == 0;
    }
    void follow(clc::stream_ref later, clc::stream_ref earlier) override
    {
        std::lock_guard<std::mutex> lock(m_);
        follows.emplace_back(later, earlier);
    }
    void drain(clc::stream_ref s) noexcept override
    {
        std::lock_guard<std::mutex> lock(m_);
        drains.push_back(s);
        busy.erase(s);
    }

    std::size_t live_bytes = 0, obtains = 0, surrenders = 0;
    std::set<void*> live;
    std::set<clc::stream_ref> busy;
    std::vector<std::pair<clc::stream_ref, clc::stream_ref>> follows;
    std::vector<clc::stream_ref> drains;

private:
    std::size_t budget_;
    std::mutex m_;
};

const std::size_t BIG = std::size_t(1) << 40;

void test_size_classes()
{
    clc::stream_pool pool(std::make_shared<fake_link>(BIG));
    CHECK(pool.class_of(0) == 256);
    CHECK(pool.class_of(1) == 256);
    CHECK(pool.class_of(256) == 256);
    CHECK(pool.class_of(257) == 512);
    CHECK(pool.class_of(300) == 512);
    CHECK(pool.class_of(std::size_t(16) << 20) == (std::size_t(16) << 20));
    CHECK(pool.class_of((std::size_t(16) << 20) + 1) == (std::size_t(17) << 20));
    CHECK(pool.class_of(std::size_t(3) << 30) == (std::size_t(3) << 30));
    bool threw = false;
    try
    {
        pool.class_of(~std::size_t(0));
    } catch (const clc::out_of_memory&)
    {
        threw = true;
    }
    CHECK(threw);
}

// The reference's allocator test sequence, on one stream.
void test_reuse_and_counters()
{
    auto link = std::make_shared<fake_link>(BIG);
    clc::stream_pool pool(link);
    const clc::stream_ref s = stream(1);
    void* a = pool.take(1000, s);
    void* b = pool.take(1000, s);
    CHECK(pool.tally().held_blocks == 2);
    pool.give(a, 1000, s);
    CHECK(pool.parked_bytes() == 1024);
    void* c = pool.take(900, s); // same class, same stream: the parked block
    CHECK(c == a);
    CHECK(pool.tally().obtained_blocks == 2);
    CHECK(pool.tally().reused == 1);
    pool.give(b, 1000, s);
    pool.give(c, 900, s);
    pool.trim();
    const auto f = pool.tally();
    CHECK(pool.parked_bytes() == 0);
    CHECK(f.held_blocks == 0 && f.held_bytes == 0);
    CHECK(f.held_peak == 2048);
    CHECK(f.obtained_bytes == 2048);
    CHECK(link->live.empty());
    CHECK(f.cross_stream == 0 && f.waits == 0 && link->follows.empty());
}

void test_owner_streams()
{
    auto link = std::make_shared<fake_link>(BIG);
    clc::stream_pool pool(link);
    const clc::stream_ref s1 = stream(1), s2 = stream(2), s3 = stream(3);

    // Parked by s1 while s1 still has work queued: s2 gets it after a
    // device-side wait on s1.
    void* a = pool.take(4096, s1);
    link->busy.insert(s1);
    pool.give(a, 4096, s1);
    void* b = pool.take(4096, s2);
    CHECK(b == a);
    CHECK(link->follows.size() == 1);
    CHECK(link->follows.size() == 1 && link->follows[0].first == s2
          && link->follows[0].second == s1);
    CHECK(pool.tally().cross_stream == 1 && pool.tally().waits == 1);
    CHECK(link->drains.empty());

    // Parked by s2 while s2 is idle: s3 takes it with no wait.
    pool.give(b, 4096, s2);
    void* c = pool.take(4096, s3);
    CHECK(c == a);
    CHECK(link->follows.size() == 1);
    CHECK(pool.tally().cross_stream == 2 && pool.tally().waits == 1);

    // A block the taker parked itself is preferred over another stream's,
    // even when the other one is older.
    void* d = pool.take(4096, s1); // a new block
    CHECK(d != c);
    link->busy.insert(s3);
    pool.give(c, 4096, s3); // older
    pool.give(d, 4096, s1); // newer
    void* e = pool.take(4096, s3);
    CHECK(e == c);
    CHECK(link->follows.size() == 1);

    // Parked with no owner: free for any stream.
    pool.give(e, 4096, clc::any_stream);
    void* f = pool.take(4096, s2);
    CHECK(f == c);
    CHECK(link->follows.size() == 1);

    // A taker that queues on no stream waits on the host for a busy owner.
    link->busy.insert(s2);
    pool.give(f, 4096, s2);
    pool.take(4096, s1); // takes d, which s1 parked
    void* g = pool.take(4096, clc::any_stream);
    CHECK(g == c);
    CHECK(link->drains.size() == 1 && link->drains[0] == s2);
    CHECK(link->follows.size() == 1);
    pool.give(g, 4096, clc::any_stream);
    pool.give(d, 4096, s1);
    pool.trim();
    CHECK(link->live.empty());
}

void test_retire()
{
    auto link = std::make_shared<fake_link>(BIG);
    clc::stream_pool pool(link);
    const clc::stream_ref s1 = stream(1), s2 = stream(2);
    void* a = pool.take(512, s1);
    link->busy.insert(s1);
    pool.give(a, 512, s1);
    pool.retire(s1); // waits for s1 and forgets it
    CHECK(link->drains.size() == 1 && link->drains[0] == s1);
    link->busy.insert(s1); // s1's handle could now name a new stream
    void* b = pool.take(512, s2);
    CHECK(b == a);
    CHECK(link->follows.empty()); // nothing ties the block to s1 any more
    CHECK(pool.tally().cross_stream == 0);
    pool.retire(s2); // nothing parked by s2: no wait
    CHECK(link->drains.size() == 1);
    pool.give(b, 512, s2);
}

void test_out_of_memory_retry()
{
    // Room for exactly two 1 MiB blocks.
    auto link = std::make_shared<fake_link>(std::size_t(2) << 20);
    clc::stream_pool pool(link);
    const clc::stream_ref s = stream(7);
    void* a = pool.take(1 << 20, s);
    void* b = pool.take(1 << 20, s);
    link->busy.insert(s);
    pool.give(a, 1 << 20, s);
    pool.give(b, 1 << 20, s);
    CHECK(pool.parked_bytes() == (std::size_t(2) << 20));
    // Another class does not fit beside the parked blocks: the pool returns
    // them (after waiting for their owner) and asks once more.
    void* c = pool.take(3 << 19, s);
    CHECK(c != nullptr);
    CHECK(pool.parked_bytes() == 0);
    CHECK(!link->drains.empty() && link->drains.back() == s);
    CHECK(pool.tally().held_blocks == 1);
    // Nothing parked and still too big: the failure stands.
    bool threw = false;
    try
    {
        pool.take(std::size_t(1) << 21, s);
    } catch (const clc::out_of_memory&)
    {
        threw = true;
    }
    CHECK(threw);
    CHECK(pool.tally().held_blocks == 1);
    pool.give(c, 3 << 19, s);
}

void test_cache_cap_evicts_oldest()
{
    auto link = std::make_shared<fake_link>(BIG);
    clc::stream_pool::limits l;
    l.cache_cap = 4096;
    clc::stream_pool pool(link, l);
    const clc::stream_ref s = stream(1);
    void* p[3];
    for (auto& q : p) q = pool.take(2048, s);
    for (auto& q : p) pool.give(q, 2048, s);
    CHECK(pool.parked_bytes() == 4096);
    CHECK(link->surrenders == 1);
    CHECK(link->live.count(p[0]) == 0); // the least recently parked went
    CHECK(link->live.count(p[1]) == 1 && link->live.count(p[2]) == 1);
    // A block larger than the whole cache goes straight back.
    void* big = pool.take(8192, s);
    pool.give(big, 8192, s);
    CHECK(link->surrenders == 2);
    CHECK(pool.parked_bytes() == 4096);
    CHECK(pool.tally().held_blocks == 2);
    // Evicting for a mid-size block keeps the rest in park order.
    void* m = pool.take(4096, s);
    pool.give(m, 4096, s);
    CHECK(pool.parked_bytes() == 4096);
    CHECK(link->live.count(p[1]) == 0 && link->live.count(p[2]) == 0);
    CHECK(pool.take(4096, s) == m);
    pool.give(m, 4096, s);
}

void test_temp_buffer_and_threads()
{
    auto link = std::make_shared<fake_link>(BIG);
    auto pool = std::make_shared<clc::stream_pool>(link);
    {
        clc::temp_buffer t(pool, 100, stream(1));
        CHECK(t.get() != nullptr && t.size() == 100);
        clc::temp_buffer z(pool, 0, stream(1));
        CHECK(z.get() == nullptr);
    }
    CHECK(pool->parked_bytes() == 256);
    std::vector<std::thread> ts;
    for (int t = 0; t < 8; ++t)
        ts.emplace_back([&, t] {
            for (int i = 0; i < 2000; ++i)
            {
                clc::temp_buffer b(pool, std::size_t(256) << (i % 6),
                                   stream(1 + (t + i) % 3));
                if (i % 97 == 0) pool->retire(stream(1 + t % 3));
            }
        });
    for (auto& t : ts) t.join();
    const auto f = pool->tally();
    CHECK(f.held_bytes == f.parked_bytes);
    pool->trim();
    CHECK(link->live.empty());
}

} // namespace

int main()
{
    test_size_classes();
    test_reuse_and_counters();
    test_owner_streams();
    test_retire();
    test_out_of_memory_retry();
    test_cache_cap_evicts_oldest();
    test_temp_buffer_and_threads();
    if (failures)
    {
        std::printf("FAILED %d of %d\n", failures, checks);
        return 1;
    }
    std::printf("ok %d\n", checks);
    return 0;
}
