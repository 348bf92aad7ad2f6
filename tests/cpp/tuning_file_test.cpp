// CPU test of the tuned-table file format (opencl-sdk_amd/csrc/tuning_file.cpp):
// round trip; tables written by another device, runtime or format version
// are ignored; files cut short, edited, padded or garbled are rejected;
// failed writes leave nothing behind; concurrent writers and readers never
// produce a mixed table (temporary file + rename). Prints "ok N" on success.
// argv[1]: an empty scratch directory.
#include "../../opencl-sdk_amd/csrc/tuning_file.hpp"

#include <dirent.h>

#include <atomic>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
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

std::string slurp(const std::string& p)
{
    std::ifstream f(p);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void spit(const std::string& p, const std::string& text)
{
    std::ofstream f(p, std::ios::trunc);
    f << text;
}

std::vector<std::string> entries(const std::string& dir)
{
    std::vector<std::string> out;
    if (DIR* d = opendir(dir.c_str()))
    {
        while (dirent* e = readdir(d))
        {
            const std::string n = e->d_name;
            if (n != "." && n != "..") out.push_back(n);
        }
        closedir(d);
    }
    return out;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const std::string id = "gfx950 cu256 hip70200000";
    const std::string p = dir + "/clc_segsort_t5_kv_gfx950.txt";
    const std::vector<int> t = { 0, 1, 1, 2, 3, 3, 4, 7, 7, -1, 123456789 };

    // Round trip, and nothing but the table's file is left.
    CHECK(clc::write_tuning_file(p, id, t));
    std::vector<int> got;
    CHECK(clc::read_tuning_file(p, id, got) && got == t);
    CHECK(entries(dir).size() == 1);
    const std::string good = slurp(p);

    // Another device, runtime or format: ignored, out untouched.
    got = { 42 };
    CHECK(!clc::read_tuning_file(p, "gfx942 cu304 hip70200000", got) && got.size() == 1);
    CHECK(!clc::read_tuning_file(p, "gfx950 cu256 hip60400000", got));
    spit(p, "clc-tuning-v1 " + good.substr(good.find(' ') + 1));
    CHECK(!clc::read_tuning_file(p, id, got));

    // Damaged copies of the good file: every one is rejected.
    const std::size_t l1 = good.find('\n'), l2 = good.find('\n', l1 + 1);
    const std::vector<std::string> bad = {
        "",                                  // empty
        good.substr(0, l1 + 1),              // header only
        good.substr(0, l2 + 1),              // no end line
        good.substr(0, good.size() - 5),     // end line cut
        good.substr(0, l1 + 4) + good.substr(l1 + 6), // a value edited away
        [&] { std::string s = good; s[l1 + 1] = '9'; return s; }(), // a value changed
        [&] { std::string s = good; s[l1 + 3] = 'x'; return s; }(), // not a number
        good + "7 7 7\n",                    // trailing junk
        [&] { std::string s = good; s.replace(s.find("end 11"), 6, "end 12"); return s; }(),
    };
    for (std::size_t i = 0; i < bad.size(); ++i)
    {
        spit(p, bad[i]);
        got = { 42 };
        const bool read = clc::read_tuning_file(p, id, got);
        if (read) std::fprintf(stderr, "damaged copy %zu was accepted\n", i);
        CHECK(!read && got.size() == 1 && got[0] == 42);
    }
    spit(p, good);
    CHECK(clc::read_tuning_file(p, id, got) && got == t);
    CHECK(!clc::read_tuning_file(dir + "/missing.txt", id, got));

    // A write that cannot happen fails quietly and leaves no temporary file.
    CHECK(!clc::write_tuning_file(dir + "/no/such/dir/x.txt", id, t));
    CHECK(!clc::write_tuning_file(p, id, {}));
    CHECK(entries(dir).size() == 1);

    // Concurrent writers of two tables and concurrent readers: a reader sees
    // one table or the other, never a mix or a partial file.
    const std::vector<int> a(32, 3), b(32, 7);
    std::atomic<bool> stop{ false };
    std::atomic<int> mixed{ 0 }, reads{ 0 };
    std::vector<std::thread> th;
    for (int w = 0; w < 4; ++w)
        th.emplace_back([&, w] {
            for (int i = 0; i < 300; ++i) clc::write_tuning_file(p, id, (i + w) % 2 ? a : b);
        });
    for (int r = 0; r < 4; ++r)
        th.emplace_back([&] {
            std::vector<int> v;
            while (!stop.load())
                if (clc::read_tuning_file(p, id, v))
                {
                    ++reads;
                    if (v != a && v != b && v != t) ++mixed;
                }
        });
    for (int w = 0; w < 4; ++w) th[w].join();
    stop = true;
    for (std::size_t i = 4; i < th.size(); ++i) th[i].join();
    CHECK(mixed.load() == 0);
    CHECK(reads.load() > 0);
    CHECK(entries(dir).size() == 1);

    if (failures)
    {
        std::printf("FAILED %d of %d\n", failures, checks);
        return 1;
    }
    std::printf("ok %d\n", checks);
    return 0;
}
