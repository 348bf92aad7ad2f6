// Operator and identity strings of clc_scan / clc_reduce.
//
// The reference specialises OpenCL C source with the caller's expression
// (binary_op, lib/include/CL/Compute/Algorithm.hpp:27-31, compiled by
// program_cache). Kernels here are compiled ahead of time, so the
// expression is recognised instead: the canonical ops:: forms
// (Algorithm.hpp:53-65), the OpenCL builtins max(a, b) / min(a, b) the
// reference tests use (capi.c:118), and the projections "a" / "b"
// (main.cpp:231-238). The identity is an OpenCL C constant expression:
// a literal, a limits macro (INT_MIN, UINT_MAX, INFINITY, ...), optionally
// negated, complemented or parenthesised.
#pragma once

#include <cstdint>
#include <stdexcept>

namespace clc {

enum op_code : int
{
    OP_PLUS = 0,
    OP_MULTIPLIES = 1,
    OP_BIT_AND = 2,
    OP_BIT_OR = 3,
    OP_BIT_XOR = 4,
    OP_MINIMUM = 5,
    OP_MAXIMUM = 6,
    OP_RIGHT = 7, // "b"
    OP_LEFT = 8, // "a"
};

//! Type codes, equal to clc_type.
enum type_code : int
{
    TY_I8 = 0, TY_U8, TY_I16, TY_U16, TY_I32, TY_U32, TY_I64, TY_U64, TY_F32,
    TY_F64
};

struct unsupported_error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

std::size_t type_size(int type);
bool type_is_float(int type);

//! Recognises expr; throws unsupported_error if it is not one of the forms
//! above and std::invalid_argument if the operator is undefined for the
//! type (bitwise operators on floating-point elements).
op_code parse_op(const char* expr, int type);

//! Evaluates identity and stores it as one element of type in out (8
//! bytes of storage). Throws unsupported_error for unrecognised text.
void parse_identity(const char* text, int type, void* out);

} // namespace clc
