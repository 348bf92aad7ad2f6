"""clc_scan / clc_reduce with operator expressions compiled at run time
(hipRTC), on the GPU through the C ABI.

The reference accepts any OpenCL C expression over a and b as the operator
and any constant expression as the identity (binary_op,
lib/include/CL/Compute/Algorithm.hpp:27-31; capi.c:114-128 and
main.cpp:231-238 use spelled-out forms). Expressions that are not one of
the canonical forms are compiled once per context; results are compared
bit-exactly with the oracle's sequential scan / reduce of the same operator
written in Python (integer wrap-around as C's), on sizes spanning one
partial tile to hundreds of tiles, inclusive, exclusive and in place.
Covers a non-commutative operator (affine-map composition: the order of
combination is checked), identities that are not canonical constants,
OpenCL builtins, unaligned buffers, build errors, and the program cache's
memory and disk hits."""
import ctypes
import os

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF


def affine(a, b):
    """x -> m*x + c mod 2^32 packed as c << 32 | m; a, then b."""
    am, ac, bm, bc = a & M32, a >> 32, b & M32, b >> 32
    return (((bm * ac + bc) & M32) << 32) | ((am * bm) & M32)


AFFINE_EXPR = ("((((b & 0xFFFFFFFFUL) * (a >> 32) + (b >> 32)) & 0xFFFFFFFFUL) << 32)"
               " | (((a & 0xFFFFFFFFUL) * (b & 0xFFFFFFFFUL)) & 0xFFFFFFFFUL)")

# (dtype, operator, identity, Python operator, Python identity, commutative)
CASES = {
    "u32_plus_one": (np.uint32, "a + b + 1", "UINT_MAX",
                     lambda a, b: a + b + 1, M32, True),
    "i8_masked_sum": (np.int8, "(a + b) & 0x7f", "0",
                      lambda a, b: (a + b) & 0x7F, 0, True),
    "u16_cast_or": (np.uint16, "(ushort)(a | b)", "0", lambda a, b: a | b, 0, True),
    "i32_builtin_max": (np.int32, "max(a, b) + 0", "INT_MIN", max, -2**31, True),
    "i64_wrapping_sum": (np.int64, "b + a - 0", "LONG_MAX + LONG_MIN + 1", lambda a, b: a + b, 0,
                         True),
    "u64_affine": (np.uint64, AFFINE_EXPR, "1", affine, 1, False),
    "f32_fmin": (np.float32, "fmin(a, b)", "INFINITY", min, float("inf"), True),
    "f64_fmax": (np.float64, "fmax(b, a)", "-INFINITY", max, float("-inf"), True),
}

SIZES = [1, 7, 1000, 4097, 100003]


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def inputs(dtype, n, seed):
    dt = np.dtype(dtype)
    rng = np.random.default_rng(seed)
    if dt.kind == "f":
        return rng.standard_normal(n).astype(dt) * 100
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)


def gpu_scan(ctx, x, op, ident, exclusive=False, inplace=False):
    import torch
    xd = dev(x)
    out = xd if inplace else torch.empty_like(xd)
    ctx.scan(xd, out, op, ident, exclusive, type_code=TYPE_OF[x.dtype])
    return host(out, x.dtype)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("n", SIZES)
def test_scan_expression_vs_oracle(ctx, case, n):
    dtype, op, ident, fn, pid, _ = CASES[case]
    say("rtc scan", case, n)
    x = inputs(dtype, n, seed=n + 17)
    for exclusive in (False, True):
        got = gpu_scan(ctx, x, op, ident, exclusive)
        want = oracle.scan_fn(x, fn, pid, exclusive)
        np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8),
                                      err_msg=f"exclusive={exclusive}")
    got = gpu_scan(ctx, x, op, ident, False, inplace=True)
    np.testing.assert_array_equal(got.view(np.uint8),
                                  oracle.scan_fn(x, fn, pid).view(np.uint8))


@pytest.mark.parametrize("case", [c for c in CASES if CASES[c][5]])
@pytest.mark.parametrize("n", [0] + SIZES)
def test_reduce_expression_vs_oracle(ctx, case, n):
    dtype, op, ident, fn, pid, _ = CASES[case]
    say("rtc reduce", case, n)
    x = inputs(dtype, n, seed=n + 29)
    init = x[0] if n else np.dtype(dtype).type(5)
    got = ctx.reduce(dev(x) if n else dev(np.zeros(1, dtype)), op, ident, init,
                     type_code=TYPE_OF[np.dtype(dtype)], n=n)
    want = oracle.reduce_fn(x, fn, pid, init.item())
    assert np.array([got], dtype=dtype).tobytes() == np.array([want], dtype=dtype).tobytes()


def test_non_commutative_order_across_many_tiles(ctx):
    """Affine-map composition over ~600 look-back tiles: a scan that combined
    any two partial results in the wrong order would differ."""
    n = 1_200_007
    say("rtc scan affine", n)
    x = inputs(np.uint64, n, seed=3)
    got = gpu_scan(ctx, x, AFFINE_EXPR, "1")
    # Closed form per element: m_i = prod m (mod 2^32); c_i by the recurrence
    # c_i = m_i' * c_{i-1} + c_i' (numpy, in uint64 with masking).
    m = (x & np.uint64(M32)).astype(np.uint64)
    c = (x >> np.uint64(32)).astype(np.uint64)
    mm = np.multiply.accumulate(m.astype(np.uint32), dtype=np.uint32).astype(np.uint64)
    np.testing.assert_array_equal(got & np.uint64(M32), mm)
    cc = np.empty(n, dtype=np.uint64)
    acc = 0
    mlist, clist = m.tolist(), c.tolist()
    for i in range(n):
        acc = (mlist[i] * acc + clist[i]) & M32
        cc[i] = acc
    np.testing.assert_array_equal(got >> np.uint64(32), cc)


def test_spelled_out_plus_matches_canonical_at_2p26(ctx):
    """'b + a + 0' runs the run-time compiled kernel, '(a + b)' the
    ahead-of-time one: bit-identical on 2^26 u32, inclusive and exclusive,
    and reduce."""
    import torch
    n = 1 << 26
    say("rtc vs canonical u32 2^26")
    x = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda")
    for ex in (False, True):
        a = torch.empty_like(x)
        b = torch.empty_like(x)
        ctx.scan(x, a, "(a + b)", "0", ex, type_code=clc_hip.U32)
        ctx.scan(x, b, "b + a + 0", "0u", ex, type_code=clc_hip.U32)
        torch.cuda.synchronize()
        assert bool(torch.equal(a, b))
    r1 = ctx.reduce(x, "(a + b)", "0", 7, type_code=clc_hip.U32)
    r2 = ctx.reduce(x, "b + a + 0", "0u", 7, type_code=clc_hip.U32)
    assert r1 == r2


def test_unaligned_buffers(ctx):
    import torch
    n = 50001
    say("rtc scan unaligned", n)
    x = inputs(np.uint32, n + 1, seed=8)
    xd = dev(x)
    out = torch.zeros_like(xd)
    L = clc_hip.lib()
    clc_hip.check(L.clc_scan(ctx.handle, ctypes.c_void_p(xd[1:].data_ptr()),
                             ctypes.c_void_p(out[1:].data_ptr()), n, clc_hip.U32,
                             b"a + b + 1", b"UINT_MAX", 0))
    want = oracle.scan_fn(x[1:], lambda a, b: a + b + 1, M32)
    np.testing.assert_array_equal(host(out, np.uint32)[1:], want)


def test_build_error_is_reported(ctx):
    import torch
    say("rtc build error")
    x = torch.zeros(16, dtype=torch.int32, device="cuda")
    L = clc_hip.lib()
    assert L.clc_scan(ctx.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                      16, clc_hip.I32, b"a +* nonsense", b"0", 0) == clc_hip.ERROR_BUILD
    assert b"build log" in L.clc_last_error()
    r = ctypes.c_int32(0)
    init = ctypes.c_int32(0)
    assert L.clc_reduce(ctx.handle, ctypes.c_void_p(x.data_ptr()), 16, clc_hip.F32,
                        b"a & b", b"0", ctypes.byref(init), ctypes.byref(r)) \
        == clc_hip.ERROR_BUILD


def test_program_cache_memory_hits():
    say("rtc program cache memory hits")
    c = clc_hip.Context()
    try:
        x = inputs(np.int32, 3000, seed=1)
        s0 = c.program_stats()
        assert (s0.builds, s0.memory_hits) == (0, 0)
        gpu_scan(c, x, "max(a, b) - 0", "INT_MIN")
        gpu_scan(c, x, "max(a, b) - 0", "INT_MIN", exclusive=True)  # same program
        gpu_scan(c, x, "max(a, b) - 0", "INT_MIN + 1")  # another identity
        c.reduce(dev(x), "max(a, b) - 0", "INT_MIN", 0)  # the reduce program
        s = c.program_stats()
        assert (s.builds, s.memory_hits, s.disk_hits) == (3, 1, 0)
    finally:
        c.close()


def test_program_cache_on_disk(tmp_path):
    import torch
    say("rtc program cache on disk", tmp_path)
    d = str(tmp_path)
    x = inputs(np.uint32, 5000, seed=2)
    want = oracle.scan_fn(x, lambda a, b: a + b + 1, M32)

    def run(c):
        xd = dev(x)
        out = torch.empty_like(xd)
        torch.cuda.synchronize()
        c.scan(xd, out, "a + b + 1", "UINT_MAX", type_code=clc_hip.U32)
        c.finish()
        np.testing.assert_array_equal(host(out, np.uint32), want)
        return c.program_stats()

    c1 = clc_hip.Context(cache_dir=d)
    s1 = run(c1)
    c1.close()
    assert (s1.builds, s1.disk_hits) == (1, 0)
    files = [f for f in os.listdir(d) if f.startswith("clc_rtc_") and f.endswith(".hsaco")]
    assert len(files) == 1
    c2 = clc_hip.Context(cache_dir=d)
    s2 = run(c2)
    c2.close()
    assert (s2.builds, s2.disk_hits) == (0, 1)
    # A foreign file under the key is rebuilt and replaced.
    with open(os.path.join(d, files[0]), "wb") as f:
        f.write(b"not a code object")
    c3 = clc_hip.Context(cache_dir=d)
    s3 = run(c3)
    c3.close()
    assert (s3.builds, s3.disk_hits) == (1, 0)
    with open(os.path.join(d, files[0]), "rb") as f:
        assert f.read(4) == b"\x7fELF"
