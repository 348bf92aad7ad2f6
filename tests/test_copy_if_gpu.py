"""clc_copy_if on the GPU through the C ABI: the elements satisfying a
predicate expression of x, in order, at the front of out, and their count
(reference clc_copy_if, compute_c.h:169-171; detail::copy_if,
Algorithm.cpp:529-567).

Checked bit-exactly against the oracle's restatement of the reference
(flags, exclusive scan, scatter) on the reference tests' own inputs and
predicates (main.cpp:316-329 "(x % 3u) == 0u", capi.c:130 "x > 0"), every
element type with +-0 / NaN / duplicates for floats, sizes around the tile
(16384 u32 elements), none / all kept, in place, unaligned input, 2^28 u32
against torch's boolean-mask selection, build errors and the program
cache. Elements of out past the count are left untouched."""
import ctypes

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def gpu_copy_if(ctx, x, pred, inplace=False, sentinel=0x5A):
    import torch
    xd = dev(x)
    if inplace:
        out = xd
    else:
        out = torch.empty_like(xd)
        out.view(torch.uint8).fill_(sentinel)
    count = ctx.copy_if(xd, out, pred, type_code=TYPE_OF[x.dtype], n=x.size)
    return host(out, x.dtype), count


@pytest.mark.parametrize("n", [0, 1, 7, 255, 256, 1000, 4097])
def test_reference_copy_if_rule(ctx, golden, n):
    say("copy_if golden", n)
    x = golden[f"hist_in_{n}"]
    if n == 0:
        import torch
        z = torch.zeros(1, dtype=torch.int32, device="cuda")
        assert ctx.copy_if(z, z, "(x % 3u) == 0u", type_code=clc_hip.U32, n=0) == 0
        return
    got, count = gpu_copy_if(ctx, x, "(x % 3u) == 0u")
    want, wc = oracle.copy_if(x, lambda v: v % 3 == 0)
    assert count == wc
    np.testing.assert_array_equal(got[:count], want)
    assert (got[count:].view(np.uint8) == 0x5A).all()


def test_capi_copy_if_vector(ctx, golden):
    say("copy_if capi vector")
    x = golden["capi_in"]
    got, count = gpu_copy_if(ctx, x, "x > 0")
    want, wc = oracle.copy_if(x, lambda v: v > 0)
    assert count == wc
    np.testing.assert_array_equal(got[:count], want)


TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32,
         np.int64, np.uint64, np.float32, np.float64]
SIZES = [1, 7, 4097, 16383, 16384, 16385, 100003]


def typed_case(dtype, n, seed):
    dt = np.dtype(dtype)
    rng = np.random.default_rng(seed)
    if dt.kind == "f":
        x = rng.standard_normal(n).astype(dt)
        x[::11] = 0.0
        x[5::13] = -0.0
        x[7::17] = np.nan
        x[3::19] = x[0]
        return x, "x < 0.25f", lambda v: v < np.float32(0.25)
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
    return x, "(x & 3) != 1", lambda v: (v & 3) != 1


@pytest.mark.parametrize("dtype", TYPES)
@pytest.mark.parametrize("n", SIZES)
def test_types_and_sizes(ctx, dtype, n):
    say("copy_if", np.dtype(dtype).name, n)
    x, pred, fn = typed_case(dtype, n, seed=n)
    got, count = gpu_copy_if(ctx, x, pred)
    want, wc = oracle.copy_if(x, fn)
    assert count == wc
    np.testing.assert_array_equal(got[:count].view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("pred,expect", [("x != x", 0), ("true", 50001), ("x >= 0u", 50001)])
def test_none_and_all_kept(ctx, pred, expect):
    say("copy_if none/all", pred)
    x = oracle.pcg32_u32(4, 4, 50001)
    got, count = gpu_copy_if(ctx, x, pred)
    assert count == expect
    np.testing.assert_array_equal(got[:count], x[:count])


@pytest.mark.parametrize("dtype", [np.uint32, np.float64, np.int8])
def test_in_place(ctx, dtype):
    n = 300001
    say("copy_if in place", np.dtype(dtype).name, n)
    x, pred, fn = typed_case(dtype, n, seed=5)
    got, count = gpu_copy_if(ctx, x, pred, inplace=True)
    want, wc = oracle.copy_if(x, fn)
    assert count == wc
    np.testing.assert_array_equal(got[:count].view(np.uint8), want.view(np.uint8))


def test_unaligned_input(ctx):
    import torch
    n = 70001
    say("copy_if unaligned", n)
    x = oracle.pcg32_u32(9, 9, n + 1)
    xd = dev(x)
    out = torch.zeros(n, dtype=torch.int32, device="cuda")
    c = ctypes.c_size_t(0)
    clc_hip.check(clc_hip.lib().clc_copy_if(ctx.handle, ctypes.c_void_p(xd[1:].data_ptr()),
                                            ctypes.c_void_p(out.data_ptr()), n, clc_hip.U32,
                                            b"(x % 5u) < 2u", ctypes.byref(c)))
    want, wc = oracle.copy_if(x[1:], lambda v: v % 5 < 2)
    assert c.value == wc
    np.testing.assert_array_equal(host(out, np.uint32)[:wc], want)


def test_full_size_2p28_against_torch(ctx):
    import torch
    n = 1 << 28
    say("copy_if 2^28 u32")
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    out = torch.empty_like(x)
    count = ctx.copy_if(x, out, "(x & 1) == 0", type_code=clc_hip.U32)
    want = x[(x & 1) == 0]
    assert count == want.numel()
    assert bool(torch.equal(out[:count], want))


def test_build_error_and_cache(ctx):
    import torch
    say("copy_if build error / cache")
    x = torch.zeros(100, dtype=torch.int32, device="cuda")
    L = clc_hip.lib()
    c = ctypes.c_size_t(0)
    assert L.clc_copy_if(ctx.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                         100, clc_hip.I32, b"x +* y", ctypes.byref(c)) == clc_hip.ERROR_BUILD
    assert b"build log" in L.clc_last_error()
    before = ctx.program_stats()
    ctx.copy_if(x, x, "x == 12345", type_code=clc_hip.I32)
    ctx.copy_if(x, x, "x == 12345", type_code=clc_hip.I32)
    after = ctx.program_stats()
    assert after.builds - before.builds == 1
    assert after.memory_hits - before.memory_hits == 1
