#include "ops.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>

namespace clc {

std::size_t type_size(int type)
{
    switch (type)
    {
        case TY_I8: case TY_U8: return 1;
        case TY_I16: case TY_U16: return 2;
        case TY_I32: case TY_U32: case TY_F32: return 4;
        case TY_I64: case TY_U64: case TY_F64: return 8;
    }
    throw std::invalid_argument("clc: unknown element type");
}

bool type_is_float(int type) { return type == TY_F32 || type == TY_F64; }

namespace {

std::string strip_spaces(const char* s)
{
    std::string out;
    for (; *s; ++s)
        if (*s != ' ' && *s != '\t' && *s != '\n' && *s != '\r') out += *s;
    return out;
}

// True when s is "(...)" with the first parenthesis closing at the end.
bool wrapped(const std::string& s)
{
    if (s.size() < 2 || s.front() != '(' || s.back() != ')') return false;
    int depth = 0;
    for (std::size_t i = 0; i < s.size(); ++i)
    {
        if (s[i] == '(') ++depth;
        else if (s[i] == ')' && --depth == 0 && i + 1 != s.size()) return false;
    }
    return true;
}

std::string unwrap(std::string s)
{
    while (wrapped(s)) s = s.substr(1, s.size() - 2);
    return s;
}

} // namespace

op_code parse_op(const char* expr, int type)
{
    if (expr == nullptr) throw std::invalid_argument("clc: null operator");
    const std::string e = unwrap(strip_spaces(expr));
    struct form
    {
        const char* text;
        op_code op;
    };
    static const form forms[] = {
        { "a+b", OP_PLUS }, { "b+a", OP_PLUS },
        { "a*b", OP_MULTIPLIES }, { "b*a", OP_MULTIPLIES },
        { "a&b", OP_BIT_AND }, { "b&a", OP_BIT_AND },
        { "a|b", OP_BIT_OR }, { "b|a", OP_BIT_OR },
        { "a^b", OP_BIT_XOR }, { "b^a", OP_BIT_XOR },
        { "(b<a)?b:a", OP_MINIMUM }, { "min(a,b)", OP_MINIMUM },
        { "(a<b)?b:a", OP_MAXIMUM }, { "max(a,b)", OP_MAXIMUM },
        { "b", OP_RIGHT }, { "a", OP_LEFT },
    };
    for (const form& f : forms)
        if (e == f.text)
        {
            if (type_is_float(type)
                && (f.op == OP_BIT_AND || f.op == OP_BIT_OR
                    || f.op == OP_BIT_XOR))
                throw std::invalid_argument(
                    "clc: bitwise operator on floating-point elements");
            return f.op;
        }
    throw unsupported_error(std::string("clc: operator expression '") + expr
                            + "' is not one of the supported forms");
}

namespace {

// A parsed constant: an integer (with sign) or a floating-point value.
struct value
{
    bool is_float = false;
    bool is_float_suffix = false; // literal ended in f/F
    double d = 0.0;
    float f = 0.0f;
    bool neg = false; // integer: value is -(mag)
    uint64_t mag = 0; // integer magnitude, or complemented bit pattern
    bool complemented = false; // ~x: bits = ~x as a 64-bit pattern
};

bool named(const std::string& s, value& v)
{
    struct entry
    {
        const char* name;
        long double val;
    };
    static const entry table[] = {
        { "CHAR_MIN", -128.0L }, { "CHAR_MAX", 127.0L },
        { "SCHAR_MIN", -128.0L }, { "SCHAR_MAX", 127.0L },
        { "UCHAR_MAX", 255.0L },
        { "SHRT_MIN", -32768.0L }, { "SHRT_MAX", 32767.0L },
        { "USHRT_MAX", 65535.0L },
        { "INT_MIN", -2147483648.0L }, { "INT_MAX", 2147483647.0L },
        { "UINT_MAX", 4294967295.0L },
    };
    for (const entry& e : table)
        if (s == e.name)
        {
            v.neg = e.val < 0;
            v.mag = (uint64_t)(e.val < 0 ? -e.val : e.val);
            return true;
        }
    if (s == "LONG_MIN")
    {
        v.neg = true;
        v.mag = 0x8000000000000000ULL;
        return true;
    }
    if (s == "LONG_MAX")
    {
        v.mag = 0x7FFFFFFFFFFFFFFFULL;
        return true;
    }
    if (s == "ULONG_MAX")
    {
        v.mag = 0xFFFFFFFFFFFFFFFFULL;
        return true;
    }
    if (s == "INFINITY" || s == "HUGE_VALF" || s == "HUGE_VAL")
    {
        v.is_float = true;
        v.d = std::numeric_limits<double>::infinity();
        v.f = std::numeric_limits<float>::infinity();
        return true;
    }
    if (s == "FLT_MAX")
    {
        v.is_float = v.is_float_suffix = true;
        v.f = std::numeric_limits<float>::max();
        v.d = v.f;
        return true;
    }
    if (s == "DBL_MAX")
    {
        v.is_float = true;
        v.d = std::numeric_limits<double>::max();
        v.f = std::numeric_limits<float>::infinity();
        return true;
    }
    return false;
}

value parse_value(std::string s)
{
    s = unwrap(s);
    if (s.empty()) throw unsupported_error("clc: empty identity");
    if (s[0] == '-' || s[0] == '+' || s[0] == '~')
    {
        value v = parse_value(s.substr(1));
        if (s[0] == '+') return v;
        if (s[0] == '~')
        {
            if (v.is_float)
                throw unsupported_error("clc: ~ applied to a floating constant");
            // Bit pattern of ~x in two's complement 64-bit.
            uint64_t bits = v.complemented ? v.mag
                                           : (v.neg ? (0 - v.mag) : v.mag);
            value r;
            r.complemented = true;
            r.mag = ~bits;
            return r;
        }
        if (v.is_float)
        {
            v.d = -v.d;
            v.f = -v.f;
        }
        else if (v.complemented)
            v.mag = 0 - v.mag;
        else
            v.neg = !v.neg;
        return v;
    }
    value v;
    if (named(s, v)) return v;
    // Numeric literal.
    const bool is_float = s.find_first_of(".eEpP") != std::string::npos
            && !(s.size() > 1 && (s[1] == 'x' || s[1] == 'X')
                 && s.find_first_of("pP") == std::string::npos);
    std::string body = s;
    if (is_float)
    {
        if (!body.empty() && (body.back() == 'f' || body.back() == 'F'))
        {
            body.pop_back();
            v.is_float_suffix = true;
        }
        char* end = nullptr;
        errno = 0;
        v.d = std::strtod(body.c_str(), &end);
        if (end == body.c_str() || *end != '\0')
            throw unsupported_error("clc: unrecognised identity '" + s + "'");
        v.f = v.is_float_suffix ? std::strtof(body.c_str(), nullptr)
                                : static_cast<float>(v.d);
        v.is_float = true;
        return v;
    }
    while (!body.empty()
           && (body.back() == 'u' || body.back() == 'U' || body.back() == 'l'
               || body.back() == 'L'))
        body.pop_back();
    char* end = nullptr;
    errno = 0;
    v.mag = std::strtoull(body.c_str(), &end, 0);
    if (end == body.c_str() || *end != '\0' || errno == ERANGE)
        throw unsupported_error("clc: unrecognised identity '" + s + "'");
    return v;
}

template <typename T> void store_int(const value& v, void* out)
{
    T t;
    if (v.is_float)
        t = static_cast<T>(v.d);
    else
    {
        const uint64_t bits =
            v.complemented ? v.mag : (v.neg ? (0 - v.mag) : v.mag);
        t = static_cast<T>(bits); // modular conversion
    }
    std::memcpy(out, &t, sizeof t);
}

} // namespace

void parse_identity(const char* text, int type, void* out)
{
    if (text == nullptr) throw std::invalid_argument("clc: null identity");
    const value v = parse_value(strip_spaces(text));
    std::memset(out, 0, 8);
    switch (type)
    {
        case TY_I8: store_int<int8_t>(v, out); return;
        case TY_U8: store_int<uint8_t>(v, out); return;
        case TY_I16: store_int<int16_t>(v, out); return;
        case TY_U16: store_int<uint16_t>(v, out); return;
        case TY_I32: store_int<int32_t>(v, out); return;
        case TY_U32: store_int<uint32_t>(v, out); return;
        case TY_I64: store_int<int64_t>(v, out); return;
        case TY_U64: store_int<uint64_t>(v, out); return;
        case TY_F32:
        {
            // ~x is an int expression; like OpenCL C, convert its value.
            float f = v.complemented ? static_cast<float>(static_cast<int64_t>(v.mag))
                    : v.is_float ? v.f
                                 : (v.neg ? -static_cast<float>(v.mag)
                                          : static_cast<float>(v.mag));
            std::memcpy(out, &f, 4);
            return;
        }
        case TY_F64:
        {
            double d = v.complemented ? static_cast<double>(static_cast<int64_t>(v.mag))
                     : v.is_float ? (v.is_float_suffix ? double(v.f) : v.d)
                                  : (v.neg ? -static_cast<double>(v.mag)
                                           : static_cast<double>(v.mag));
            std::memcpy(out, &d, 8);
            return;
        }
    }
    throw std::invalid_argument("clc: unknown element type");
}

} // namespace clc
