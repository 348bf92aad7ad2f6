"""clc_fill, clc_sequence, clc_transform and clc_transform2 on the GPU
through the C ABI (reference compute_c.h:142-155, Algorithm.cpp:331-398).

fill is checked bit for bit on every element width; sequence against the
oracle's restatement (wrap-around for integers, rounded product then sum
for floats) and the reference test's closed form (capi.c:139-141);
transform / transform2 against numpy evaluations of the same expressions
(integer wrap, float32 rounding per operation), including the reference
test's "(float)x * 0.5f" (capi.c:124-127), type conversions, sizes around
the 2048-element tile and build errors."""
import ctypes

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say

pytestmark = pytest.mark.gpu

TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32,
         np.int64, np.uint64, np.float32, np.float64]
SIZES = [1, 7, 2047, 2048, 2049, 100003]


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def empty(n, dtype):
    import torch
    t = {1: torch.uint8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[np.dtype(dtype).itemsize]
    out = torch.empty(n + 1, dtype=t, device="cuda")
    out.view(torch.uint8).fill_(0xA5)
    return out


@pytest.mark.parametrize("dtype", TYPES)
@pytest.mark.parametrize("n", SIZES)
def test_fill(ctx, dtype, n):
    say("fill", np.dtype(dtype).name, n)
    dt = np.dtype(dtype)
    v = np.array([-3.25 if dt.kind == "f" else 0x5B], dtype=dt)
    out = empty(n, dtype)
    ctx.fill(out, v[0], type_code=TYPE_OF[dt], n=n)
    got = host(out, dt)
    assert (got[:n].view(np.uint8).reshape(n, -1) == v.view(np.uint8)).all()
    assert (got[n:].view(np.uint8) == 0xA5).all()


@pytest.mark.parametrize("dtype", TYPES)
@pytest.mark.parametrize("n", SIZES)
def test_sequence(ctx, dtype, n):
    say("sequence", np.dtype(dtype).name, n)
    dt = np.dtype(dtype)
    start, step = (0.5, 0.25) if dt.kind == "f" else (100, 7 if dt.kind == "u" else -3)
    out = empty(n, dtype)
    ctx.sequence(out, start, step, type_code=TYPE_OF[dt], n=n)
    want = oracle.sequence(n, dt, start, step)
    np.testing.assert_array_equal(host(out, dt)[:n].view(np.uint8), want.view(np.uint8))


def test_reference_capi_sequence_fill_transform(ctx, golden):
    """capi.c:124-127 and :139-145."""
    import torch
    say("capi sequence / fill / transform")
    n = golden["capi_in"].size
    dst = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.sequence(dst, 5, 2, type_code=clc_hip.I32)
    d = host(dst, np.int32)
    assert d[0] == 5 and d[-1] == 5 + 2 * (n - 1)
    ctx.fill(dst, 0, type_code=clc_hip.I32)
    assert host(dst, np.int32)[0] == 0
    x = golden["capi_in"]
    f = torch.empty(n, dtype=torch.float32, device="cuda")
    ctx.transform(dev(x), f, "(float)x * 0.5f", in_type=clc_hip.I32, out_type=clc_hip.F32)
    got = host(f, np.float32)
    assert got[3] == np.float32(x[3]) * np.float32(0.5)
    np.testing.assert_array_equal(got, x.astype(np.float32) * np.float32(0.5))


CASES = [
    # (in dtype, out dtype, expression, numpy function)
    (np.uint32, np.float32, "x * 0.18f + 32.0f",
     lambda x: (x.astype(np.float32) * np.float32(0.18)) + np.float32(32.0)),
    (np.int32, np.int32, "x * 3 - 7", lambda x: (x.astype(np.int64) * 3 - 7).astype(np.int32)),
    (np.uint8, np.uint16, "popcount(x) + (x >> 4)",
     lambda x: (np.unpackbits(x[:, None], axis=1).sum(1) + (x >> 4)).astype(np.uint16)),
    (np.float64, np.int64, "convert_long(x * 4.0)", lambda x: (x * 4.0).astype(np.int64)),
    (np.int16, np.float64, "(double)x / 3.0", lambda x: x.astype(np.float64) / 3.0),
    (np.uint64, np.uint32, "(uint)(x >> 32) ^ (uint)x",
     lambda x: ((x >> np.uint64(32)) ^ (x & np.uint64(0xFFFFFFFF))).astype(np.uint32)),
]


def inputs(dtype, n, seed):
    dt = np.dtype(dtype)
    rng = np.random.default_rng(seed)
    if dt.kind == "f":
        return (rng.standard_normal(n) * 1000).astype(dt)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("n", SIZES)
def test_transform(ctx, case, n):
    ti, to, expr, fn = CASES[case]
    say("transform", np.dtype(ti).name, np.dtype(to).name, expr, n)
    x = inputs(ti, n, seed=case * 100 + n)
    out = empty(n, to)
    ctx.transform(dev(x), out, expr, in_type=TYPE_OF[np.dtype(ti)],
                  out_type=TYPE_OF[np.dtype(to)], n=n)
    got = host(out, np.dtype(to))
    with np.errstate(over="ignore", invalid="ignore"):
        want = fn(x).astype(to)
    np.testing.assert_array_equal(got[:n].view(np.uint8), want.view(np.uint8))
    assert (got[n:].view(np.uint8) == 0xA5).all()


CASES2 = [
    (np.uint32, np.uint32, np.uint32, "x + y", lambda x, y: x + y),
    (np.int8, np.float32, np.float32, "(float)x * y",
     lambda x, y: x.astype(np.float32) * y),
    (np.int64, np.int32, np.int64, "max(x, (long)y) - min(x, (long)y)",
     lambda x, y: np.maximum(x, y.astype(np.int64)) - np.minimum(x, y.astype(np.int64))),
]


@pytest.mark.parametrize("case", range(len(CASES2)))
@pytest.mark.parametrize("n", SIZES)
def test_transform2(ctx, case, n):
    t1, t2, to, expr, fn = CASES2[case]
    say("transform2", expr, n)
    x = inputs(t1, n, seed=case + n)
    y = inputs(t2, n, seed=case + n + 1)
    out = empty(n, to)
    ctx.transform2(dev(x), dev(y), out, expr, type1=TYPE_OF[np.dtype(t1)],
                   type2=TYPE_OF[np.dtype(t2)], out_type=TYPE_OF[np.dtype(to)], n=n)
    got = host(out, np.dtype(to))
    with np.errstate(over="ignore"):
        want = fn(x, y).astype(to)
    np.testing.assert_array_equal(got[:n].view(np.uint8), want.view(np.uint8))


def test_in_place_and_errors(ctx):
    import torch
    say("transform in place / errors")
    x = inputs(np.int32, 5000, seed=1)
    xd = dev(x)
    ctx.transform(xd, xd, "x ^ 0x55", in_type=clc_hip.I32, out_type=clc_hip.I32)
    np.testing.assert_array_equal(host(xd, np.int32), x ^ 0x55)
    L = clc_hip.lib()
    assert L.clc_transform(ctx.handle, ctypes.c_void_p(xd.data_ptr()), clc_hip.I32,
                           ctypes.c_void_p(xd.data_ptr()), clc_hip.I32, 10,
                           b"x +* nonsense") == clc_hip.ERROR_BUILD
    assert b"build log" in L.clc_last_error()
    z = torch.zeros(1, dtype=torch.int32, device="cuda")
    v = ctypes.c_int32(1)
    assert L.clc_fill(ctx.handle, ctypes.c_void_p(z.data_ptr()), 0, clc_hip.I32,
                      ctypes.byref(v)) == clc_hip.SUCCESS
    assert L.clc_fill(ctx.handle, None, 5, 42, ctypes.byref(v)) == clc_hip.ERROR_INVALID_ARGUMENT
