// Tuned launch tables on disk (tuning_file.hpp).
#include "tuning_file.hpp"

#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <unistd.h>

namespace clc {
namespace {

const char* const MAGIC = "clc-tuning-v2 ";

uint64_t fnv1a(const std::string& s)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (unsigned char c : s)
    {
        h ^= c;
        h *= 0x100000001b3ull;
    }
    return h;
}

std::string hex(uint64_t v)
{
    char buf[17];
    std::snprintf(buf, sizeof buf, "%016llx", static_cast<unsigned long long>(v));
    return buf;
}

std::string values_line(const std::vector<int>& v)
{
    std::string s;
    for (std::size_t i = 0; i < v.size(); ++i)
    {
        if (i) s += ' ';
        s += std::to_string(v[i]);
    }
    return s;
}

//! Parses a line of space-separated decimal integers strictly.
bool parse_ints(const std::string& line, std::vector<int>& out)
{
    std::vector<int> v;
    std::size_t i = 0;
    while (i < line.size())
    {
        if (line[i] == ' ')
        {
            ++i;
            continue;
        }
        std::size_t j = i;
        if (line[j] == '-') ++j;
        const std::size_t digits = j;
        while (j < line.size() && line[j] >= '0' && line[j] <= '9') ++j;
        if (j == digits || j - digits > 9) return false;
        if (j < line.size() && line[j] != ' ') return false;
        v.push_back(std::atoi(line.substr(i, j - i).c_str()));
        i = j;
    }
    out.swap(v);
    return true;
}

} // namespace

bool read_tuning_file(const std::string& path, const std::string& identity,
                      std::vector<int>& out)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string header, values, end, extra;
    if (!std::getline(f, header) || !std::getline(f, values) || !std::getline(f, end))
        return false;
    if (std::getline(f, extra) && !extra.empty()) return false;
    if (header != MAGIC + identity) return false;
    std::vector<int> v;
    if (!parse_ints(values, v) || v.empty()) return false;
    const std::string want = "end " + std::to_string(v.size()) + " " + hex(fnv1a(values));
    if (end != want) return false;
    out.swap(v);
    return true;
}

bool write_tuning_file(const std::string& path, const std::string& identity,
                       const std::vector<int>& values)
{
    if (path.empty() || values.empty()) return false;
    static std::atomic<unsigned> serial{ 0 };
    const std::string tmp = path + ".tmp." + std::to_string(::getpid()) + "."
        + std::to_string(serial.fetch_add(1));
    bool ok = false;
    try
    {
        const std::string line = values_line(values);
        {
            std::ofstream f(tmp, std::ios::trunc);
            if (f)
            {
                f << MAGIC << identity << '\n'
                  << line << '\n'
                  << "end " << values.size() << ' ' << hex(fnv1a(line)) << '\n';
                f.flush();
                ok = static_cast<bool>(f);
            }
        }
        if (ok) ok = std::rename(tmp.c_str(), path.c_str()) == 0;
    } catch (...)
    {
        ok = false;
    }
    if (!ok) std::remove(tmp.c_str());
    return ok;
}

} // namespace clc
