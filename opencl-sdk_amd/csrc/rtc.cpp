#include "rtc.hpp"

#include "context.hpp"
#include "ops.hpp"

#include <hip/hiprtc.h>

#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iterator>

namespace clc {
namespace rtc {

// The kernel-body headers, embedded by the build (generated source).
extern const int embedded_header_count;
extern const char* const embedded_header_names[];
extern const char* const embedded_header_texts[];

namespace {

// Compiler options. -fwrapv: signed integer expressions wrap, as OpenCL C
// implementations do; -ffp-contract=off: every floating-point operation of
// the expression rounds on its own; no packed FP32 instructions, like the
// ahead-of-time kernels.
const char* const OPTIONS[] = {
    "-O3", "-std=c++17", "-fwrapv", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
};

const char* type_name(int type)
{
    switch (type)
    {
        case TY_I8: return "signed char";
        case TY_U8: return "unsigned char";
        case TY_I16: return "short";
        case TY_U16: return "unsigned short";
        case TY_I32: return "int";
        case TY_U32: return "unsigned int";
        case TY_I64: return "long";
        case TY_U64: return "unsigned long";
        case TY_F32: return "float";
        case TY_F64: return "double";
    }
    throw std::invalid_argument("clc: unknown element type");
}

// OpenCL C's limit macros (OpenCL C 3.0 sections 6.15.2.1, 6.15.3) and the
// scalar forms of its builtins that operator expressions use: integer
// (abs, abs_diff, add_sat, sub_sat, hadd, rhadd, mul_hi, mad24, mul24,
// rotate, popcount, clz), common (max, min, clamp, mix, sign), mad, select
// and the convert_<type> / as_<type> conversions. Math functions (fmax,
// fmin, fabs, fma, sqrt, ...) are HIP's. They live in namespace clc_ocl,
// with the expression, so they hide HIP's global max / min / abs.
const char* const PRELUDE = R"CLC(
#ifndef CHAR_BIT
#define CHAR_BIT 8
#endif
#ifndef SCHAR_MAX
#define SCHAR_MAX 127
#define SCHAR_MIN (-127 - 1)
#define CHAR_MAX SCHAR_MAX
#define CHAR_MIN SCHAR_MIN
#define UCHAR_MAX 255
#define SHRT_MAX 32767
#define SHRT_MIN (-32767 - 1)
#define USHRT_MAX 65535
#define INT_MAX 2147483647
#define INT_MIN (-2147483647 - 1)
#define UINT_MAX 0xffffffffU
#define LONG_MAX 0x7fffffffffffffffL
#define LONG_MIN (-0x7fffffffffffffffL - 1)
#define ULONG_MAX 0xffffffffffffffffUL
#endif
#ifndef FLT_MAX
#define FLT_MAX 0x1.fffffep127f
#define FLT_MIN 0x1.0p-126f
#define FLT_EPSILON 0x1.0p-23f
#define DBL_MAX 0x1.fffffffffffffp1023
#define DBL_MIN 0x1.0p-1022
#define DBL_EPSILON 0x1.0p-52
#endif
#ifndef MAXFLOAT
#define MAXFLOAT FLT_MAX
#endif
#ifndef INFINITY
#define INFINITY __builtin_huge_valf()
#endif
#ifndef HUGE_VALF
#define HUGE_VALF __builtin_huge_valf()
#endif
#ifndef HUGE_VAL
#define HUGE_VAL __builtin_huge_val()
#endif
#ifndef NAN
#define NAN __builtin_nanf("")
#endif

namespace clc_ocl {

typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
typedef unsigned long ulong;

template <typename X, typename Y>
using common_t = typename std::common_type<X, Y>::type;
template <typename X>
using unsigned_t = typename std::make_unsigned<X>::type;
template <typename X>
using if_int = typename std::enable_if<std::is_integral<X>::value, X>::type;
template <typename X>
using if_float = typename std::enable_if<std::is_floating_point<X>::value, X>::type;
template <typename X> constexpr int bits_of() { return int(sizeof(X) * 8); }

template <typename X, typename Y>
__device__ __forceinline__ common_t<X, Y> max(X x, Y y)
{
    using C = common_t<X, Y>;
    return C(x) < C(y) ? C(y) : C(x);
}
template <typename X, typename Y>
__device__ __forceinline__ common_t<X, Y> min(X x, Y y)
{
    using C = common_t<X, Y>;
    return C(y) < C(x) ? C(y) : C(x);
}
template <typename X, typename L, typename H>
__device__ __forceinline__ X clamp(X x, L lo, H hi)
{
    return min(max(x, X(lo)), X(hi));
}
template <typename X>
__device__ __forceinline__ unsigned_t<if_int<X>> abs(X x)
{
    using U = unsigned_t<X>;
    return x < X(0) ? U(U(0) - U(x)) : U(x);
}
template <typename X> __device__ __forceinline__ if_float<X> abs(X x)
{
    return __builtin_fabs(x);
}
template <typename X, typename Y>
__device__ __forceinline__ unsigned_t<if_int<common_t<X, Y>>> abs_diff(X x, Y y)
{
    using C = common_t<X, Y>;
    using U = unsigned_t<C>;
    return C(x) > C(y) ? U(U(C(x)) - U(C(y))) : U(U(C(y)) - U(C(x)));
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> add_sat(X x, Y y)
{
    using U = unsigned_t<X>;
    X r;
    if (!__builtin_add_overflow(x, X(y), &r)) return r;
    if constexpr (std::is_signed<X>::value)
        return X(y) > X(0) ? X(~(U(1) << (bits_of<X>() - 1))) : X(U(1) << (bits_of<X>() - 1));
    else
        return X(~U(0));
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> sub_sat(X x, Y y)
{
    using U = unsigned_t<X>;
    X r;
    if (!__builtin_sub_overflow(x, X(y), &r)) return r;
    if constexpr (std::is_signed<X>::value)
        return X(y) < X(0) ? X(~(U(1) << (bits_of<X>() - 1))) : X(U(1) << (bits_of<X>() - 1));
    else
        return X(0);
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> hadd(X x, Y y)
{
    const X b = X(y);
    return X((x >> 1) + (b >> 1) + (x & b & 1));
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> rhadd(X x, Y y)
{
    const X b = X(y);
    return X((x >> 1) + (b >> 1) + ((x | b) & 1));
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> mul_hi(X x, Y y)
{
    if constexpr (sizeof(X) == 8)
    {
        if constexpr (std::is_signed<X>::value)
            return X((__int128(x) * __int128(X(y))) >> 64);
        else
            return X((static_cast<unsigned __int128>(x) * X(y)) >> 64);
    }
    else
    {
        using W = typename std::conditional<std::is_signed<X>::value, long, unsigned long>::type;
        return X((W(x) * W(X(y))) >> bits_of<X>());
    }
}
template <typename X, typename Y, typename Z>
__device__ __forceinline__ if_int<X> mad_hi(X x, Y y, Z z)
{
    return X(mul_hi(x, y) + X(z));
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> mul24(X x, Y y)
{
    return X(x * X(y));
}
template <typename X, typename Y, typename Z>
__device__ __forceinline__ if_int<X> mad24(X x, Y y, Z z)
{
    return X(x * X(y) + X(z));
}
template <typename X, typename Y> __device__ __forceinline__ if_int<X> rotate(X x, Y y)
{
    using U = unsigned_t<X>;
    const int s = int(U(y) % bits_of<X>());
    const U u = U(x);
    return s == 0 ? x : X(U(u << s) | U(u >> (bits_of<X>() - s)));
}
template <typename X> __device__ __forceinline__ if_int<X> popcount(X x)
{
    return X(__builtin_popcountll((unsigned long long)unsigned_t<X>(x)));
}
template <typename X> __device__ __forceinline__ if_int<X> clz(X x)
{
    const unsigned long long u = (unsigned long long)unsigned_t<X>(x);
    return u == 0 ? X(bits_of<X>()) : X(__builtin_clzll(u) - (64 - bits_of<X>()));
}
template <typename X, typename Y, typename Z>
__device__ __forceinline__ common_t<X, Y> mad(X x, Y y, Z z)
{
    using C = common_t<X, Y>;
    return C(C(x) * C(y) + C(z));
}
template <typename X, typename Y, typename A>
__device__ __forceinline__ common_t<X, Y> mix(X x, Y y, A a)
{
    using C = common_t<X, Y>;
    return C(C(x) + (C(y) - C(x)) * C(a));
}
template <typename X> __device__ __forceinline__ if_float<X> sign(X x)
{
    return x > X(0) ? X(1) : (x < X(0) ? X(-1) : (x == x ? x : X(0)));
}
template <typename A, typename B, typename C>
__device__ __forceinline__ common_t<A, B> select(A a, B b, C c)
{
    return c ? common_t<A, B>(b) : common_t<A, B>(a);
}

#define CLC_OCL_CONVERSIONS(name, type)                                       \
    template <typename X> __device__ __forceinline__ type convert_##name(X x) \
    {                                                                         \
        return static_cast<type>(x);                                          \
    }                                                                         \
    template <typename X> __device__ __forceinline__ type as_##name(X x)      \
    {                                                                         \
        static_assert(sizeof(X) == sizeof(type), "as_" #name ": size");       \
        type r;                                                               \
        __builtin_memcpy(&r, &x, sizeof r);                                   \
        return r;                                                             \
    }
CLC_OCL_CONVERSIONS(char, signed char)
CLC_OCL_CONVERSIONS(uchar, unsigned char)
CLC_OCL_CONVERSIONS(short, short)
CLC_OCL_CONVERSIONS(ushort, unsigned short)
CLC_OCL_CONVERSIONS(int, int)
CLC_OCL_CONVERSIONS(uint, unsigned int)
CLC_OCL_CONVERSIONS(long, long)
CLC_OCL_CONVERSIONS(ulong, unsigned long)
CLC_OCL_CONVERSIONS(float, float)
CLC_OCL_CONVERSIONS(double, double)
#undef CLC_OCL_CONVERSIONS

} // namespace clc_ocl
)CLC";

uint64_t fnv1a(const std::string& data, uint64_t h = 14695981039346656037ull)
{
    for (unsigned char c : data)
    {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string hex(uint64_t v)
{
    char text[17];
    std::snprintf(text, sizeof text, "%016llx", static_cast<unsigned long long>(v));
    return text;
}

struct program_guard
{
    hiprtcProgram p = nullptr;
    ~program_guard()
    {
        if (p) (void)hiprtcDestroyProgram(&p);
    }
};

void rtc_check(hiprtcResult r, const char* what)
{
    if (r != HIPRTC_SUCCESS)
        throw std::runtime_error(std::string("clc: ") + what + " failed: "
                                 + hiprtcGetErrorString(r));
}

} // namespace

std::string base_arch(const std::string& arch)
{
    const std::size_t colon = arch.find(':');
    return colon == std::string::npos ? arch : arch.substr(0, colon);
}

namespace {

// Headers and the builtins prelude, leaving namespace clc_ocl open.
std::string source_open()
{
    std::string s;
    s += "#include <cstdint>\n#include <type_traits>\n";
    s += "#include \"scan_kernels.hpp\"\n#include \"elementwise_kernels.hpp\"\n";
    s += PRELUDE;
    s += "namespace clc_ocl {\n";
    return s;
}

// The same with "typedef <element type> T;".
std::string source_head(int type)
{
    return source_open() + "typedef " + type_name(type) + " T;\n";
}

// "return (OUT)(<expression>);" with the expression on lines of its own,
// named name in the build log.
std::string return_expression(const char* out_type, const std::string& expr,
                              const char* name)
{
    return std::string("        return (") + out_type + ")(\n#line 1 \"" + name + "\"\n"
        + expr + "\n#line 1 \"clc_rtc_kernels\"\n);\n";
}

} // namespace

std::string transform_source(int in_type, int in2_type, int out_type,
                             const std::string& expression)
{
    // As the reference's transform kernels: T_IN x (T_IN1 x, T_IN2 y) in,
    // (T_OUT)(EXPR) out (Algorithm.cpp:30-56).
    const bool two = in2_type >= 0;
    std::string s = source_open();
    s += std::string("typedef ") + type_name(out_type) + " T_OUT;\n";
    if (two)
        s += std::string("typedef ") + type_name(in_type) + " T_IN1;\ntypedef "
            + type_name(in2_type) + " T_IN2;\n"
            + "struct user_expr\n{\n"
              "    static __device__ __forceinline__ T_OUT apply(T_IN1 x, T_IN2 y)\n    {\n"
            + return_expression("T_OUT", expression, "expression") + "    }\n};\n";
    else
        s += std::string("typedef ") + type_name(in_type) + " T_IN;\n"
            + "struct user_expr\n{\n"
              "    static __device__ __forceinline__ T_OUT apply(T_IN x)\n    {\n"
            + return_expression("T_OUT", expression, "expression") + "    }\n};\n";
    s += "} // namespace clc_ocl\n";
    if (two)
        s += "extern \"C\" __global__\n"
             "__launch_bounds__(clc::dev::EW_BLOCK) void clc_rtc_transform2(\n"
             "    const clc_ocl::T_IN1* in1, const clc_ocl::T_IN2* in2, clc_ocl::T_OUT* out,\n"
             "    size_t n)\n{\n"
             "    clc::dev::transform2_tile<clc_ocl::T_IN1, clc_ocl::T_IN2, clc_ocl::T_OUT,\n"
             "                              clc_ocl::user_expr, clc::dev::EW_BLOCK,\n"
             "                              clc::dev::EW_ITEMS>(in1, in2, out, n);\n}\n";
    else
        s += "extern \"C\" __global__\n"
             "__launch_bounds__(clc::dev::EW_BLOCK) void clc_rtc_transform(\n"
             "    const clc_ocl::T_IN* in, clc_ocl::T_OUT* out, size_t n)\n{\n"
             "    clc::dev::transform_tile<clc_ocl::T_IN, clc_ocl::T_OUT, clc_ocl::user_expr,\n"
             "                             clc::dev::EW_BLOCK, clc::dev::EW_ITEMS>(\n"
             "        in, out, n);\n}\n";
    return s;
}

std::string predicate_source(int type, const std::string& predicate)
{
    std::string s = source_head(type);
    // As the reference's select kernels: flag = (PRED) ? 1 : 0 with the
    // element named x (Algorithm.cpp:215-237).
    s += "struct user_pred\n{\n"
         "    static __device__ __forceinline__ bool test(T x)\n"
         "    {\n"
         "        return (\n#line 1 \"predicate\"\n";
    s += predicate;
    s += "\n#line 1 \"clc_rtc_kernels\"\n) ? true : false;\n    }\n};\n"
         "} // namespace clc_ocl\n";
    s += "extern \"C\" __global__\n"
         "__launch_bounds__(clc::dev::SCAN_BLOCK) void clc_rtc_copy_if(\n"
         "    const clc_ocl::T* in, clc_ocl::T* out, size_t n, int vec_ok,\n"
         "    uint64_t* status, uint32_t* tile_counter, uint32_t* count)\n{\n"
         "    clc::dev::select_tile<clc_ocl::T, clc_ocl::user_pred, clc::dev::SCAN_BLOCK,\n"
         "                          clc::dev::scan_items<clc_ocl::T>()>(\n"
         "        in, out, n, vec_ok, status, tile_counter, count);\n}\n";
    return s;
}

std::string operator_source(int type, const std::string& op,
                            const std::string& identity, int kernels)
{
    std::string s = source_head(type);
    // The expressions sit on lines of their own so the build log points at
    // them by name ("operator:1:7: error: ...").
    s += "struct user_op\n{\n"
         "    static __device__ __forceinline__ T apply(T a, T b)\n"
         "    {\n"
         "        return (T)(\n#line 1 \"operator\"\n";
    s += op;
    s += "\n#line 1 \"clc_rtc_kernels\"\n);\n    }\n};\n";
    s += "__device__ __forceinline__ T user_identity()\n{\n    return (T)(\n"
         "#line 1 \"identity\"\n";
    s += identity;
    s += "\n#line 1 \"clc_rtc_kernels\"\n);\n}\n} // namespace clc_ocl\n";
    if (kernels & WITH_SCAN)
        s += "extern \"C\" __global__\n"
             "__launch_bounds__(clc::dev::SCAN_BLOCK) void clc_rtc_scan(\n"
             "    const clc_ocl::T* in, clc_ocl::T* out, size_t n, int exclusive,\n"
             "    int vec_ok, uint64_t* status, uint32_t* tile_counter)\n{\n"
             "    clc::dev::scan_tile<clc_ocl::T, clc_ocl::user_op, clc::dev::SCAN_BLOCK,\n"
             "                        clc::dev::scan_items<clc_ocl::T>()>(\n"
             "        in, out, n, clc_ocl::user_identity(), exclusive, vec_ok,\n"
             "        status, tile_counter);\n}\n";
    if (kernels & WITH_REDUCE)
        s += "extern \"C\" __global__\n"
             "__launch_bounds__(clc::dev::REDUCE_BLOCK) void clc_rtc_reduce(\n"
             "    const clc_ocl::T* in, size_t n, int vec_ok, clc_ocl::T* out,\n"
             "    int apply_init, clc_ocl::T init)\n{\n"
             "    clc::dev::reduce_partial<clc_ocl::T, clc_ocl::user_op,\n"
             "                             clc::dev::REDUCE_BLOCK>(\n"
             "        in, n, clc_ocl::user_identity(), vec_ok, out, apply_init, init);\n}\n";
    return s;
}

std::vector<char> compile(const std::string& source, const std::string& arch)
{
    program_guard g;
    rtc_check(hiprtcCreateProgram(&g.p, source.c_str(), "clc_rtc_kernels",
                                  embedded_header_count, embedded_header_texts,
                                  embedded_header_names),
              "hiprtcCreateProgram");
    const std::string target = "--offload-arch=" + base_arch(arch);
    std::vector<const char*> opts{ target.c_str() };
    for (const char* o : OPTIONS) opts.push_back(o);
    const hiprtcResult r =
        hiprtcCompileProgram(g.p, static_cast<int>(opts.size()), opts.data());
    if (r != HIPRTC_SUCCESS)
    {
        std::size_t n = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(g.p, &n) == HIPRTC_SUCCESS && n > 0)
        {
            log.resize(n);
            if (hiprtcGetProgramLog(g.p, &log[0]) != HIPRTC_SUCCESS) log.clear();
            while (!log.empty() && log.back() == '\0') log.pop_back();
        }
        throw build_error(std::string("clc: build failed (")
                          + hiprtcGetErrorString(r) + ")\n--- build log for "
                          + base_arch(arch) + " ---\n" + log);
    }
    std::size_t size = 0;
    rtc_check(hiprtcGetCodeSize(g.p, &size), "hiprtcGetCodeSize");
    std::vector<char> code(size);
    rtc_check(hiprtcGetCode(g.p, code.data()), "hiprtcGetCode");
    return code;
}

// ---- program cache -------------------------------------------------------------

program_cache::program_cache(std::string disk_dir, std::string arch)
    : disk_dir_(std::move(disk_dir)), arch_(base_arch(arch))
{
    int major = 0, minor = 0, runtime = 0;
    (void)hiprtcVersion(&major, &minor);
    (void)hipRuntimeGetVersion(&runtime);
    identity_ = "clc-rtc-v1\n" + arch_ + "\nhiprtc " + std::to_string(major) + "."
        + std::to_string(minor) + "\nhip " + std::to_string(runtime) + "\n";
    for (const char* o : OPTIONS) identity_ += std::string(o) + "\n";
    // The embedded headers are part of every source.
    for (int i = 0; i < embedded_header_count; ++i)
        identity_ += std::string(embedded_header_names[i]) + "\n"
            + embedded_header_texts[i] + "\n";
}

program_cache::~program_cache()
{
    for (auto& kv : programs_)
        if (kv.second.module) (void)hipModuleUnload(kv.second.module);
}

std::string program_cache::key_for(const std::string& source) const
{
    uint64_t h = fnv1a(source);
    h = fnv1a(std::string(1, '\0') + identity_, h);
    return hex(h);
}

bool program_cache::load_from_disk(const std::string& key, hipModule_t& module)
{
    if (disk_dir_.empty()) return false;
    std::ifstream f(disk_dir_ + "/clc_rtc_" + key + ".hsaco", std::ios::binary);
    if (!f) return false;
    std::vector<char> code((std::istreambuf_iterator<char>(f)),
                           std::istreambuf_iterator<char>());
    // A code object is an ELF file; anything else is stale or foreign.
    if (code.size() < 4 || code[0] != 0x7f || code[1] != 'E' || code[2] != 'L'
        || code[3] != 'F')
        return false;
    return hipModuleLoadData(&module, code.data()) == hipSuccess;
}

void program_cache::store_to_disk(const std::string& key,
                                  const std::vector<char>& code) const
{
    if (disk_dir_.empty()) return;
    const std::string path = disk_dir_ + "/clc_rtc_" + key + ".hsaco";
    const std::string tmp = path + ".tmp." + std::to_string(::getpid());
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return;
        f.write(code.data(), static_cast<std::streamsize>(code.size()));
        if (!f) return;
    }
    // Persisting is best effort: a failed rename leaves no partial file.
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

hipFunction_t program_cache::function(const std::string& source, const char* name)
{
    const std::string key = key_for(source);
    auto it = programs_.find(key);
    if (it != programs_.end())
        ++stats_.memory_hits;
    else
    {
        program p;
        if (load_from_disk(key, p.module))
            ++stats_.disk_hits;
        else
        {
            const std::vector<char> code = compile(source, arch_);
            CLC_HIP(hipModuleLoadData(&p.module, code.data()));
            store_to_disk(key, code);
            ++stats_.builds;
        }
        it = programs_.emplace(key, std::move(p)).first;
    }
    program& p = it->second;
    auto f = p.functions.find(name);
    if (f == p.functions.end())
    {
        hipFunction_t fn = nullptr;
        CLC_HIP(hipModuleGetFunction(&fn, p.module, name));
        f = p.functions.emplace(name, fn).first;
    }
    return f->second;
}

} // namespace rtc
} // namespace clc
