"""C-ABI checks that need no GPU: the library loads, exports every function
include/compute_c.h declares, and reports errors the way the reference's
C API does (lib/src/Compute/CApi.cpp:60-115, lib/test/Compute/capi.c)."""
import ctypes
import os
import shutil
import subprocess
import re

import pytest

import clc_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(clc_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    L = clc_hip.lib()
    declared = declared_functions()
    assert len(declared) >= 18
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # The binding covers every declared function.
    assert declared == set(clc_hip.SIGNATURES)


def test_null_arguments_are_invalid_argument():
    L = clc_hip.lib()
    assert L.clc_device_count(None) == clc_hip.ERROR_INVALID_ARGUMENT
    assert "null" in L.clc_last_error().decode()
    assert L.clc_context_create(0, None) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_finish(None) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_sort(None, None, clc_hip.U32, None, 10, 0) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_scan(None, None, None, 10, clc_hip.U32, b"(a + b)", b"0", 0) \
        == clc_hip.ERROR_INVALID_ARGUMENT
    x = ctypes.c_uint32(0)
    assert L.clc_reduce(None, None, 10, clc_hip.U32, b"(a + b)", b"0",
                        ctypes.byref(x), ctypes.byref(x)) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_histogram_even(None, None, 10, clc_hip.U32, None, 4, 0.0, 1.0) \
        == clc_hip.ERROR_INVALID_ARGUMENT
    L.clc_context_release(None)  # no-op, like the reference


def test_last_error_cleared_on_success():
    L = clc_hip.lib()
    assert L.clc_operator_supported(clc_hip.U32, b"nonsense", b"0") == clc_hip.ERROR_BUILD
    assert len(L.clc_last_error()) > 0
    assert L.clc_operator_supported(clc_hip.U32, b"(a + b)", b"0") == clc_hip.SUCCESS
    assert L.clc_last_error() == b""
    assert L.clc_last_cl_error() == 0


def test_device_enumeration_without_gpu():
    L = clc_hip.lib()
    count = ctypes.c_size_t(123)
    assert L.clc_device_count(ctypes.byref(count)) == clc_hip.SUCCESS
    info = clc_hip.DeviceInfo()
    assert L.clc_device_get_info(count.value, ctypes.byref(info)) \
        == clc_hip.ERROR_INVALID_ARGUMENT
    if count.value == 0:
        h = ctypes.c_void_p()
        assert L.clc_context_create(clc_hip.DEFAULT_DEVICE, ctypes.byref(h)) \
            == clc_hip.ERROR_NO_DEVICE
        assert h.value is None


@pytest.mark.parametrize("t,op,ident", [
    (clc_hip.U32, "(a + b)", "0"), (clc_hip.I32, "a+b", "0"),
    (clc_hip.U32, "(a * b)", "1"), (clc_hip.U32, "(a & b)", "(~0)"),
    (clc_hip.U64, "(a | b)", "0"), (clc_hip.I8, "(a ^ b)", "0"),
    (clc_hip.U32, "((b < a) ? b : a)", "UINT_MAX"),
    (clc_hip.I32, "((a < b) ? b : a)", "INT_MIN"),
    (clc_hip.I32, "max(a, b)", "INT_MIN"), (clc_hip.F32, "min(a, b)", "INFINITY"),
    (clc_hip.F32, "((a < b) ? b : a)", "(-INFINITY)"),
    (clc_hip.F64, "(a + b)", "0.0"), (clc_hip.I32, "b", "0"),
    (clc_hip.I64, "((b < a) ? b : a)", "LONG_MAX"),
    (clc_hip.U64, "((a < b) ? b : a)", "0"), (clc_hip.F32, "(a*b)", "1.0f"),
])
def test_canonical_operators_supported(t, op, ident):
    assert clc_hip.lib().clc_operator_supported(t, op.encode(), ident.encode()) \
        == clc_hip.SUCCESS, clc_hip.lib().clc_last_error()


@pytest.mark.parametrize("t,op,ident,status", [
    (clc_hip.U32, "a - b", "0", clc_hip.SUCCESS),  # compiles; associativity is the caller's
    (clc_hip.U32, "x +* nonsense", "0", clc_hip.ERROR_BUILD),
    (clc_hip.F32, "(a & b)", "0", clc_hip.ERROR_BUILD),  # as OpenCL C: no & on float
    (clc_hip.U32, "(a + b)", "zero", clc_hip.ERROR_BUILD),
    (clc_hip.F32, "(a + b)", "1 + 2", clc_hip.SUCCESS),
    (clc_hip.U32, "(a + b", "0", clc_hip.ERROR_BUILD),
    (42, "(a + b)", "0", clc_hip.ERROR_INVALID_ARGUMENT),
    (42, "a - b", "0", clc_hip.ERROR_INVALID_ARGUMENT),
])
def test_operator_expressions_compile_or_fail(t, op, ident, status):
    """Non-canonical expressions are compiled at run time (no device needed
    for the check); one that does not compile is CLC_ERROR_BUILD with the
    build log, as the reference's capi test expects (capi.c:151-156)."""
    L = clc_hip.lib()
    assert L.clc_operator_supported(t, op.encode(), ident.encode()) == status, \
        L.clc_last_error().decode()
    if status == clc_hip.ERROR_BUILD:
        assert b"build log" in L.clc_last_error()


def test_build_log_names_the_expression():
    L = clc_hip.lib()
    assert L.clc_operator_supported(clc_hip.I32, b"a + undefined_name", b"0") \
        == clc_hip.ERROR_BUILD
    err = L.clc_last_error().decode()
    assert "operator:1:" in err and "undefined_name" in err
    assert L.clc_operator_supported(clc_hip.I32, b"a + b", b"bad_identity") \
        == clc_hip.ERROR_BUILD
    assert "identity:1:" in L.clc_last_error().decode()


# OpenCL C scalar builtins, limit macros, conversions and type names that
# operator expressions may use, on every element type they apply to.
INT_TYPES = [clc_hip.I8, clc_hip.U8, clc_hip.I16, clc_hip.U16, clc_hip.I32,
             clc_hip.U32, clc_hip.I64, clc_hip.U64]
FLOAT_TYPES = [clc_hip.F32, clc_hip.F64]


@pytest.mark.parametrize("t", INT_TYPES)
@pytest.mark.parametrize("op,ident", [
    ("max(a, b)", "0"), ("min(a, b) + clamp(a, 0, 100)", "0"),
    ("abs(a) + abs_diff(a, b) + add_sat(a, b) + sub_sat(a, b)", "0"),
    ("hadd(a, b) ^ rhadd(a, b) ^ mul_hi(a, b) ^ mad_hi(a, b, a)", "0"),
    ("rotate(a, 3) + popcount(b) + clz(a) + mad24(a, b, a) + mul24(a, b)", "0"),
    ("select(a, b, a < b) + mad(a, b, 1)", "CHAR_MAX + SHRT_MIN + UINT_MAX + LONG_MIN"),
    ("(T)(convert_int(a) + convert_ulong(b) + as_uint(1.0f) + (uint)(uchar)a)", "INT_MIN"),
])
def test_integer_builtins_compile(t, op, ident):
    L = clc_hip.lib()
    assert L.clc_operator_supported(t, op.encode(), ident.encode()) == clc_hip.SUCCESS, \
        L.clc_last_error().decode()[-2000:]


@pytest.mark.parametrize("t", FLOAT_TYPES)
@pytest.mark.parametrize("op,ident", [
    ("fmax(a, b)", "-INFINITY"), ("fmin(a, b)", "INFINITY"),
    ("max(a, b) + min(a, b) + clamp(a, -1.0f, 1.0f)", "-FLT_MAX"),
    ("mad(a, b, 1) + mix(a, b, 0.5f) + sign(a) + select(a, b, a < b)", "0.0"),
    ("fabs(a) + abs(b) + sqrt(a) + fma(a, b, a) + convert_float(b)", "DBL_MAX"),
    ("a + b", "NAN"), ("a * b", "HUGE_VALF + FLT_MIN + FLT_EPSILON + MAXFLOAT"),
])
def test_float_builtins_compile(t, op, ident):
    L = clc_hip.lib()
    assert L.clc_operator_supported(t, op.encode(), ident.encode()) == clc_hip.SUCCESS, \
        L.clc_last_error().decode()[-2000:]


@pytest.mark.parametrize("t,pred,status", [
    (clc_hip.U32, "(x % 3u) == 0u", clc_hip.SUCCESS),  # main.cpp:321
    (clc_hip.I32, "x > 0", clc_hip.SUCCESS),  # capi.c:130
    (clc_hip.F32, "x < 0.5f && !isnan(x)", clc_hip.SUCCESS),
    (clc_hip.U8, "x & 1", clc_hip.SUCCESS),
    (clc_hip.F64, "fabs(x) > 1.0", clc_hip.SUCCESS),
    (clc_hip.I64, "popcount(x) > 32", clc_hip.SUCCESS),
    (clc_hip.U32, "x +* y", clc_hip.ERROR_BUILD),
    (clc_hip.F32, "x & 1", clc_hip.ERROR_BUILD),
    (42, "x > 0", clc_hip.ERROR_INVALID_ARGUMENT),
])
def test_copy_if_predicates_compile_or_fail(t, pred, status):
    L = clc_hip.lib()
    assert L.clc_predicate_supported(t, pred.encode()) == status, \
        L.clc_last_error().decode()[-2000:]
    if status == clc_hip.ERROR_BUILD:
        assert b"build log" in L.clc_last_error() and b"predicate:1:" in L.clc_last_error()


def test_copy_if_null_arguments():
    L = clc_hip.lib()
    c = ctypes.c_size_t(7)
    assert L.clc_copy_if(None, None, None, 10, clc_hip.U32, b"x > 0", ctypes.byref(c)) \
        == clc_hip.ERROR_INVALID_ARGUMENT


@pytest.mark.parametrize("ti,t2,to,expr,status", [
    (clc_hip.I32, -1, clc_hip.F32, "(float)x * 0.5f", clc_hip.SUCCESS),  # capi.c:124
    (clc_hip.U32, -1, clc_hip.F32, "x * 0.18f + 32.0f", clc_hip.SUCCESS),  # compute-pipeline
    (clc_hip.U32, clc_hip.U32, clc_hip.U32, "x + y", clc_hip.SUCCESS),  # compute-pipeline
    (clc_hip.F64, clc_hip.I8, clc_hip.I64, "convert_long(x) * y + (T_OUT)sizeof(T_IN2)",
     clc_hip.SUCCESS),
    (clc_hip.U8, -1, clc_hip.U16, "popcount(x) + clz(x)", clc_hip.SUCCESS),
    (clc_hip.I32, -1, clc_hip.I32, "x +* nonsense", clc_hip.ERROR_BUILD),
    (clc_hip.I32, -1, clc_hip.I32, "x + y", clc_hip.ERROR_BUILD),  # no y in transform
    (clc_hip.I32, 42, clc_hip.I32, "x + y", clc_hip.ERROR_INVALID_ARGUMENT),
])
def test_transform_expressions_compile_or_fail(ti, t2, to, expr, status):
    L = clc_hip.lib()
    assert L.clc_transform_supported(ti, t2, to, expr.encode()) == status, \
        L.clc_last_error().decode()[-2000:]
    if status == clc_hip.ERROR_BUILD:
        assert b"expression:1:" in L.clc_last_error()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_header_and_c_test_compile_as_c99(tmp_path):
    """include/compute_c.h is a C header: the C end-to-end test compiles
    with -std=c99 -pedantic -Werror and links against the library."""
    exe = str(tmp_path / "capi_test")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-O1",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "capi_test.c"), "-o", exe,
                    "-L", os.path.join(ROOT, "opencl-sdk_amd"), "-lclc_hip", "-lm"],
                   check=True)
    assert os.path.exists(exe)


def test_buffer_handles_without_device():
    L = clc_hip.lib()
    b = ctypes.c_void_p(123)
    assert L.clc_buffer_alloc(None, 16, ctypes.byref(b)) == clc_hip.ERROR_INVALID_ARGUMENT
    assert b.value is None
    assert L.clc_buffer_size(None) == 0
    assert L.clc_buffer_data(None) is None
    L.clc_buffer_release(None)
    assert L.clc_buffer_read(None, None, 0, 0, None) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_buffer_copy(None, None, 0, None, 0, 0) == clc_hip.ERROR_INVALID_ARGUMENT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cxx_header_compiles_and_links(tmp_path):
    """include/compute.hpp (the C++ algorithm API over the C ABI) compiles
    with warnings as errors, and the C++ GPU test built on it links against
    the library (running it needs a GPU)."""
    from test_compute_api_gpu import compile_cmd
    exe = str(tmp_path / "compute_api_test")
    r = subprocess.run(compile_cmd(os.path.join(ROOT, "tests", "cpp", "compute_api_test.cpp"),
                                   exe), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(exe)
