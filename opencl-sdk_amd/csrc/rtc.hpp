// Run-time specialisation of clc_scan / clc_reduce for operator and
// identity expressions the ahead-of-time kernels do not cover.
//
// The reference pastes the caller's OP and IDENTITY text into OpenCL C
// sources and builds them through a program cache keyed by the expanded
// source, the build options and the device, in memory and on disk
// (lib/src/Compute/ProgramCache.cpp:142-260, Algorithm.cpp:280-290,
// :405-415). Here the expression is pasted into a C++ functor that
// instantiates the same scan / reduce kernel bodies the library ships
// (scan_kernels.hpp, embedded in the library at build time), hipRTC compiles
// the result for the device's gfx architecture, and a per-context
// program_cache keeps the loaded modules (memory hits), the code objects in
// the context's cache directory (disk hits), and counts both and the builds.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace clc {

//! A run-time compilation failed (CLC_ERROR_BUILD at the C boundary);
//! what() carries the compiler's build log.
struct build_error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

namespace rtc {

//! Kernel names of the generated sources.
constexpr const char* SCAN_KERNEL = "clc_rtc_scan";
constexpr const char* REDUCE_KERNEL = "clc_rtc_reduce";

//! Kernels a generated source holds (bit set).
enum kernel_set : int
{
    WITH_SCAN = 1,
    WITH_REDUCE = 2
};

//! Source of the scan and/or reduce kernel (kernels: a kernel_set) over
//! elements of type (a type_code) with the operator expression op over a
//! and b and the identity expression identity, both in OpenCL C scalar
//! syntax:
//!   scan:   clc_rtc_scan(const T* in, T* out, size_t n, int exclusive,
//!                        int vec_ok, uint64_t* status, uint32_t* counter)
//!   reduce: clc_rtc_reduce(const T* in, size_t n, int vec_ok, T* out,
//!                          int apply_init, T init)
std::string operator_source(int type, const std::string& op,
                            const std::string& identity, int kernels);

//! Source of the copy_if kernel over elements of type with the predicate
//! expression predicate over x (OpenCL C scalar syntax):
//!   clc_rtc_copy_if(const T* in, T* out, size_t n, int vec_ok,
//!                   uint64_t* status, uint32_t* counter, uint32_t* count)
std::string predicate_source(int type, const std::string& predicate);
constexpr const char* COPY_IF_KERNEL = "clc_rtc_copy_if";

//! Source of the transform kernel (in2_type < 0: out[i] = (T_OUT)(expr)
//! with x = in[i]) or the transform2 kernel (x = in1[i], y = in2[i]):
//!   clc_rtc_transform(const T_IN* in, T_OUT* out, size_t n)
//!   clc_rtc_transform2(const T_IN1* in1, const T_IN2* in2, T_OUT* out, size_t n)
std::string transform_source(int in_type, int in2_type, int out_type,
                             const std::string& expression);
constexpr const char* TRANSFORM_KERNEL = "clc_rtc_transform";
constexpr const char* TRANSFORM2_KERNEL = "clc_rtc_transform2";

//! Compiles source with hipRTC for arch (a gfx target name such as
//! gfx950); returns the code object. Throws build_error with the log.
std::vector<char> compile(const std::string& source, const std::string& arch);

//! The gfx target of a device architecture string ("gfx950:sramecc+:xnack-"
//! gives "gfx950").
std::string base_arch(const std::string& arch);

struct cache_stats
{
    std::size_t memory_hits = 0;
    std::size_t disk_hits = 0;
    std::size_t builds = 0;
};

//! Loaded run-time programs of one context (one device). Not thread-safe,
//! like the context that owns it.
class program_cache {
public:
    //! disk_dir: directory for code objects ("" for memory only); arch:
    //! the device's architecture string.
    program_cache(std::string disk_dir, std::string arch);
    ~program_cache();
    program_cache(const program_cache&) = delete;
    program_cache& operator=(const program_cache&) = delete;

    //! Kernel name of the program built from source, building or loading
    //! it first when needed. The context's device must be current.
    hipFunction_t function(const std::string& source, const char* name);

    const cache_stats& stats() const { return stats_; }

    //! Cache key of source on this device: hex FNV-1a over the source, the
    //! compiler options and the compiler / runtime versions.
    std::string key_for(const std::string& source) const;

private:
    struct program
    {
        hipModule_t module = nullptr;
        std::map<std::string, hipFunction_t> functions;
    };
    bool load_from_disk(const std::string& key, hipModule_t& module);
    void store_to_disk(const std::string& key, const std::vector<char>& code) const;

    std::string disk_dir_;
    std::string arch_;
    std::string identity_; // options + versions folded into every key
    std::map<std::string, program> programs_;
    cache_stats stats_;
};

} // namespace rtc
} // namespace clc
