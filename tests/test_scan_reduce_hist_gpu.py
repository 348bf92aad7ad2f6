"""Scan, reduce and histogram_even parity on the GPU, through the C ABI.

Mirrors the reference's test_scan / test_reduce /
test_copy_if_and_histogram (lib/test/Compute/main.cpp:182-244, :314-337)
and the C-ABI checks (capi.c:114-128), then every type/operator pair
against the oracle. Integer results are bit-exact. Floating-point sums are
made bit-exact by using integer-valued inputs whose partial sums are exactly
representable; a random-float sum is checked with a relative tolerance of
1e-5 of the running sum of |x| (order of combination differs)."""
import ctypes

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 7, 255, 256, 1000, 4097]

OP_TEXT = {"plus": ("(a + b)", None), "multiplies": ("(a * b)", None),
           "bit_and": ("(a & b)", None), "bit_or": ("(a | b)", None),
           "bit_xor": ("(a ^ b)", None), "minimum": ("((b < a) ? b : a)", None),
           "maximum": ("((a < b) ? b : a)", None), "right": ("b", None)}


def identity_of(op, dtype):
    dt = np.dtype(dtype)
    if op in ("plus", "bit_or", "bit_xor", "right"):
        return "0", 0
    if op == "multiplies":
        return "1", 1
    if op == "bit_and":
        return "(~0)", -1 if dt.kind == "i" else np.iinfo(dt).max
    if dt.kind == "f":
        return ("INFINITY", np.inf) if op == "minimum" else ("(-INFINITY)", -np.inf)
    info = np.iinfo(dt)
    return (str(info.max), info.max) if op == "minimum" else (str(info.min), info.min)


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def gpu_scan(ctx, x, op="(a + b)", ident="0", exclusive=False, inplace=False):
    import torch
    xd = dev(x)
    out = xd if inplace else torch.empty_like(xd)
    if x.size:
        ctx.scan(xd, out, op, ident, exclusive, type_code=TYPE_OF[x.dtype])
    else:
        ctx.scan(xd, out, op, ident, exclusive, type_code=TYPE_OF[x.dtype], n=0)
    return host(out, x.dtype)


# ---- scan ------------------------------------------------------------------

@pytest.mark.parametrize("n", SIZES)
def test_reference_scan_vectors(ctx, golden, n):
    say("scan golden", n)
    x = golden[f"scan_in_{n}"]
    np.testing.assert_array_equal(gpu_scan(ctx, x), golden[f"scan_expect_inclusive_{n}"])
    np.testing.assert_array_equal(gpu_scan(ctx, x, exclusive=True),
                                  golden[f"scan_expect_exclusive_{n}"])
    np.testing.assert_array_equal(gpu_scan(ctx, x, inplace=True),
                                  golden[f"scan_expect_inclusive_{n}"])


def test_reference_projection_and_max_scan(ctx, golden):
    say("scan projection / max")
    v = golden["proj_in"]
    np.testing.assert_array_equal(gpu_scan(ctx, v, "b", "0"), v)
    np.testing.assert_array_equal(gpu_scan(ctx, v, "((a < b) ? b : a)", "INT_MIN"),
                                  golden["max_scan_expect"])


def test_capi_scan_vector(ctx, golden):
    say("scan capi vector")
    x = golden["capi_in"]
    out = gpu_scan(ctx, x, "(a + b)", "0")
    s, _ = golden["capi_expect_sum_max"]
    assert out[0] == x[0] and out[-1] == np.int32(s)


def test_c1_exclusive_scan_2p20(ctx):
    n = 1 << 20
    say("scan C1 exclusive 2^20 of pcg32 & 0xFF")
    x = oracle.pcg32_u32(42, 55, n) & np.uint32(0xFF)
    np.testing.assert_array_equal(gpu_scan(ctx, x, exclusive=True),
                                  oracle.scan(x, "plus", 0, exclusive=True))


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32,
                                   np.uint32, np.int64, np.uint64, np.float32,
                                   np.float64])
@pytest.mark.parametrize("op", list(OP_TEXT))
@pytest.mark.parametrize("exclusive", [False, True])
def test_scan_types_ops_vs_oracle(ctx, dtype, op, exclusive):
    dt = np.dtype(dtype)
    if dt.kind == "f" and op.startswith("bit_"):
        pytest.skip("bitwise operators are undefined for floats")
    n = 70001
    say("scan", dt.name, op, "exclusive" if exclusive else "inclusive", n)
    rng = np.random.default_rng(17)
    if dt.kind == "f":
        # Small integer values: every partial sum/product is exact in any order.
        hi = 2 if op == "multiplies" else 8
        x = rng.integers(0, hi, size=n).astype(dt)
        if op == "multiplies":
            x[rng.random(n) < 0.99] = 1
    else:
        info = np.iinfo(dt)
        x = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
    text, _ = OP_TEXT[op]
    ident_text, ident = identity_of(op, dt)
    got = gpu_scan(ctx, x, text, ident_text, exclusive)
    want = oracle.scan(x, op, ident, exclusive)
    np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))


def test_scan_float_sum_tolerance(ctx):
    n = 1 << 20
    say("scan f32 random sum", n)
    x = np.random.default_rng(4).standard_normal(n).astype(np.float32)
    got = gpu_scan(ctx, x)
    want = np.cumsum(x.astype(np.float64))
    scale = np.cumsum(np.abs(x.astype(np.float64)))
    assert np.all(np.abs(got - want) <= 1e-5 * scale + 1e-6)


def test_scan_full_size_property(ctx):
    import torch
    n = 1 << 28
    say("scan u32 2^28 difference property")
    x = torch.randint(0, 256, (n,), dtype=torch.int32, device="cuda")
    out = torch.empty_like(x)
    ctx.scan(x, out, "(a + b)", "0", exclusive=True, type_code=clc_hip.U32)
    torch.cuda.synchronize()
    # u32 arithmetic wraps modulo 2^32; compare the differences modulo 2^32.
    mask = 0xFFFFFFFF
    o = out.to(torch.int64) & mask
    first_ok = int(o[0]) == 0
    diffs_ok = bool((((o[1:] - o[:-1]) & mask) == x[:-1].to(torch.int64)).all())
    total_ok = (int(o[-1]) + int(x[-1])) & mask == int(x.to(torch.int64).sum()) & mask
    assert first_ok and diffs_ok and total_ok


def test_scan_unaligned(ctx):
    import torch
    n = 50001
    say("scan unaligned", n)
    x = oracle.pcg32_u32(1, 2, n + 1)
    xd = dev(x)
    out = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    clc_hip.check(clc_hip.lib().clc_scan(ctx.handle, ctypes.c_void_p(xd[1:].data_ptr()),
                                         ctypes.c_void_p(out[1:].data_ptr()), n,
                                         clc_hip.U32, b"(a + b)", b"0", 0))
    np.testing.assert_array_equal(host(out, np.uint32)[1:], oracle.scan(x[1:], "plus", 0))


# ---- reduce -------------------------------------------------------------------

@pytest.mark.parametrize("n", SIZES)
def test_reference_reduce_vectors(ctx, golden, n):
    say("reduce golden", n)
    x = golden[f"reduce_in_{n}"]
    e = golden[f"reduce_expect_{n}"]
    xd = dev(x) if n else dev(np.zeros(1, np.uint32))
    assert ctx.reduce(xd, "(a + b)", "0", 7, type_code=clc_hip.U32, n=n) == e[0]
    if n:
        assert ctx.reduce(xd, "((a < b) ? b : a)", "0", 0, type_code=clc_hip.U32) == e[1]
        assert ctx.reduce(xd, "((b < a) ? b : a)", "UINT_MAX", 0xFFFFFFFF,
                          type_code=clc_hip.U32) == e[2]


def test_reference_float_minmax_and_capi(ctx, golden):
    say("reduce floats and capi vector")
    f = (np.arange(10000) % 17 - 8).astype(np.float32)
    fd = dev(f)
    assert ctx.reduce(fd, "((a < b) ? b : a)", "(-INFINITY)", -np.inf) == 8.0
    assert ctx.reduce(fd, "((b < a) ? b : a)", "INFINITY", np.inf) == -8.0
    x = golden["capi_in"]
    s, m = golden["capi_expect_sum_max"]
    xd = dev(x)
    assert ctx.reduce(xd, "(a + b)", "0", 0) == np.int32(s)
    assert ctx.reduce(xd, "max(a, b)", "INT_MIN", np.iinfo(np.int32).min) == m


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32,
                                   np.uint32, np.int64, np.uint64, np.float32,
                                   np.float64])
@pytest.mark.parametrize("op", ["plus", "multiplies", "bit_and", "bit_or", "bit_xor",
                                "minimum", "maximum"])
@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_reduce_types_ops_vs_oracle(ctx, dtype, op, n):
    dt = np.dtype(dtype)
    if dt.kind == "f" and op.startswith("bit_"):
        pytest.skip("bitwise operators are undefined for floats")
    say("reduce", dt.name, op, n)
    rng = np.random.default_rng(n)
    if dt.kind == "f":
        x = rng.integers(0, 8, size=n).astype(dt)
        if op == "multiplies":
            x = np.where(rng.random(n) < 0.001, 2, 1).astype(dt)
    else:
        info = np.iinfo(dt)
        x = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
    text, _ = OP_TEXT[op]
    ident_text, ident = identity_of(op, dt)
    init = ident
    got = ctx.reduce(dev(x), text, ident_text, init, type_code=TYPE_OF[dt])
    want = oracle.reduce(x, op, ident, init)
    assert np.array([got], dtype=dt).view(np.uint8).tolist() == \
        np.array([want], dtype=dt).view(np.uint8).tolist()


def test_reduce_empty_gives_init(ctx):
    say("reduce empty")
    x = dev(np.zeros(4, np.uint32))
    assert ctx.reduce(x, "(a + b)", "0", 42, type_code=clc_hip.U32, n=0) == 42


# ---- histogram_even -------------------------------------------------------------

def gpu_hist(ctx, x, bins, lower, upper):
    import torch
    h = torch.full((max(bins, 1),), -1, dtype=torch.int32, device="cuda")
    xd = dev(x) if x.size else dev(np.zeros(1, x.dtype))
    ctx.histogram_even(xd, h[:bins] if bins else h[:0], lower, upper,
                       type_code=TYPE_OF[x.dtype], n=x.size)
    return host(h, np.uint32)[:bins]


@pytest.mark.parametrize("n", SIZES)
def test_reference_histogram_vectors(ctx, golden, n):
    say("histogram golden", n)
    x = golden[f"hist_in_{n}"]
    np.testing.assert_array_equal(gpu_hist(ctx, x, 10, 0.0, 1000.0),
                                  golden[f"hist_expect_{n}"])


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32,
                                   np.uint32, np.int64, np.uint64, np.float32,
                                   np.float64])
@pytest.mark.parametrize("bins", [1, 10, 256, 5000, 100000])
def test_histogram_types_vs_oracle(ctx, dtype, bins):
    dt = np.dtype(dtype)
    n = 300001
    say("histogram", dt.name, bins, n)
    rng = np.random.default_rng(bins)
    if dt.kind == "f":
        x = (rng.standard_normal(n) * 40).astype(dt)
        lo, up = -100.25, 90.5
    else:
        info = np.iinfo(dt)
        lo_i, hi_i = max(info.min, -20000), min(info.max, 40000)
        x = rng.integers(lo_i, hi_i, size=n, endpoint=True).astype(dt)
        lo, up = float(lo_i) + 3.0, float(hi_i) - 7.0
    got = gpu_hist(ctx, x, bins, lo, up)
    np.testing.assert_array_equal(got, oracle.histogram_even(x, bins, lo, up))


def test_histogram_empty_and_errors(ctx):
    say("histogram empty / errors")
    np.testing.assert_array_equal(gpu_hist(ctx, np.zeros(0, np.uint32), 7, 0.0, 1.0),
                                  np.zeros(7, np.uint32))
    x = dev(np.zeros(4, np.float32))
    import torch
    h = torch.zeros(4, dtype=torch.int32, device="cuda")
    assert clc_hip.lib().clc_histogram_even(ctx.handle, ctypes.c_void_p(x.data_ptr()), 4,
                                            clc_hip.F32, ctypes.c_void_p(h.data_ptr()),
                                            4, 1.0, 1.0) == clc_hip.ERROR_INVALID_ARGUMENT
