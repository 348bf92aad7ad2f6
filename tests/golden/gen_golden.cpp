// Writes the inputs and expected outputs of the reference's own Compute tests
// (lib/test/Compute/main.cpp, lib/test/Compute/capi.c) as raw little-endian
// files, for make_golden.py to pack into reference_tests.npz.
//
// Inputs are drawn exactly as main.cpp:45-52 draws them (std::mt19937 and
// std::uniform_int_distribution under libstdc++, the standard library the
// reference's Linux CI uses). Expected outputs come from the same library
// routines the reference tests compare against: std::stable_sort
// (main.cpp:259-263), std::sort (main.cpp:293, :308), std::accumulate /
// running sums (main.cpp:188, :216-221), x / 100 bin counting (main.cpp:330).
// Nothing here runs the reference library itself.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <string>
#include <vector>

static std::string outdir;

template <typename T>
static void dump(const std::string& name, const std::vector<T>& v)
{
    const std::string path = outdir + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f)
    {
        std::perror(path.c_str());
        std::exit(1);
    }
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

// main.cpp:45-52
static std::vector<uint32_t> random_uints(std::size_t n, uint32_t max,
                                          unsigned seed)
{
    std::mt19937 rng(seed);
    std::uniform_int_distribution<uint32_t> dist(0, max);
    std::vector<uint32_t> v(n);
    for (auto& x : v) x = dist(rng);
    return v;
}

int main(int argc, char** argv)
{
    if (argc != 2)
    {
        std::fprintf(stderr, "usage: gen_golden OUTDIR\n");
        return 2;
    }
    outdir = argv[1];
    // main.cpp:54, without the two largest sizes (kept out of the fixture
    // for size; the GPU tests cover them against the oracle).
    const std::size_t sizes[] = { 0, 1, 7, 255, 256, 1000, 4097 };

    for (std::size_t n : sizes)
    {
        const std::string s = std::to_string(n);
        // test_sort, main.cpp:248-271
        auto keys = random_uints(n, 0xFFFFFFFFu, unsigned(n) + 2);
        std::vector<std::pair<uint32_t, uint32_t>> ref(n);
        for (std::size_t i = 0; i < n; ++i) ref[i] = { keys[i], uint32_t(i) };
        std::stable_sort(ref.begin(), ref.end(),
                         [](const std::pair<uint32_t, uint32_t>& a,
                            const std::pair<uint32_t, uint32_t>& b) {
                             return a.first < b.first;
                         });
        std::vector<uint32_t> rk(n), rv(n);
        for (std::size_t i = 0; i < n; ++i)
        {
            rk[i] = ref[i].first;
            rv[i] = ref[i].second;
        }
        dump("sort_keys_" + s, keys);
        dump("sort_expect_keys_" + s, rk);
        dump("sort_expect_values_" + s, rv);

        // test_scan, main.cpp:209-229
        auto sc = random_uints(n, 100, unsigned(n) + 1);
        std::vector<uint32_t> inc(n), exc(n);
        uint32_t run = 0;
        for (std::size_t i = 0; i < n; ++i)
        {
            exc[i] = run;
            run += sc[i];
            inc[i] = run;
        }
        dump("scan_in_" + s, sc);
        dump("scan_expect_inclusive_" + s, inc);
        dump("scan_expect_exclusive_" + s, exc);

        // test_reduce, main.cpp:184-197
        auto rd = random_uints(n, 1000, unsigned(n));
        std::vector<uint32_t> red(3);
        red[0] = std::accumulate(rd.begin(), rd.end(), uint32_t(7));
        red[1] = n ? *std::max_element(rd.begin(), rd.end()) : 0u;
        red[2] = n ? *std::min_element(rd.begin(), rd.end()) : 0xFFFFFFFFu;
        dump("reduce_in_" + s, rd);
        dump("reduce_expect_" + s, red);

        // test_copy_if_and_histogram, main.cpp:316-335
        auto hi = random_uints(n, 1000, unsigned(n) + 3);
        std::vector<uint32_t> href(10, 0);
        for (uint32_t x : hi)
            if (x < 1000) ++href[x / 100];
        dump("hist_in_" + s, hi);
        dump("hist_expect_" + s, href);
    }

    // Stability with duplicates, main.cpp:273-287.
    {
        auto dup = random_uints(50000, 15, 99);
        std::vector<uint32_t> idx(dup.size());
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t b) { return dup[a] < dup[b]; });
        std::vector<uint8_t> dup8(dup.begin(), dup.end());
        dump("stable_keys_u8", dup8);
        dump("stable_expect_values", idx);
    }

    // Signed extremes, main.cpp:289-295.
    {
        std::vector<int32_t> ints{ 5, -3, 0, INT32_MIN, 7, -1, INT32_MAX, 2 };
        auto si = ints;
        std::sort(si.begin(), si.end());
        dump("ints_in", ints);
        dump("ints_expect", si);
    }

    // Float keys, descending, with -0, +inf, -inf, main.cpp:297-311.
    {
        std::vector<float> floats(3000);
        std::mt19937 rng(5);
        std::uniform_real_distribution<float> dist(-1e6f, 1e6f);
        for (auto& f : floats) f = dist(rng);
        floats[0] = -0.0f;
        floats[1] = __builtin_inff();
        floats[2] = -__builtin_inff();
        auto sf = floats;
        std::sort(sf.begin(), sf.end(), [](float a, float b) { return a > b; });
        dump("floats_in", floats);
        dump("floats_expect_desc", sf);
    }

    // Non-commutative and max scans, main.cpp:231-243.
    {
        std::vector<int32_t> v(5000);
        for (std::size_t i = 0; i < v.size(); ++i) v[i] = int32_t(i * 7 % 101);
        std::vector<int32_t> mx(v.size());
        int32_t m = INT32_MIN;
        for (std::size_t i = 0; i < v.size(); ++i) mx[i] = m = std::max(m, v[i]);
        dump("proj_in", v);
        dump("max_scan_expect", mx);
    }

    // C-ABI test vector, capi.c:104-128: N = 10007 ints (i*7919 % 1000) - 500.
    {
        const std::size_t N = 10007;
        std::vector<int32_t> host(N);
        int64_t expected = 0;
        for (std::size_t i = 0; i < N; ++i)
        {
            host[i] = int32_t((i * 7919u) % 1000u) - 500;
            expected += host[i];
        }
        std::vector<int64_t> e{ expected, 499 };
        dump("capi_in", host);
        dump("capi_expect_sum_max", e);
    }
    return 0;
}
