"""clc_scan_noalloc, the scan that allocates nothing, through the C ABI.

Its look-back scratch is a static per-GPU device array instead of a block
of the device pool. Each call is compared bit for bit with clc_scan (the
same kernels over pool scratch) on the same input, two calls in a row with
the second on a smaller input (the array keeps the first call's contents,
so the second proves it is cleared), for several element types and
operators; the pool's counters must not move during the allocation-free
calls. Also: expressions are refused (compiling one allocates), n == 0 is
a no-op, null buffers are rejected."""
import ctypes

import numpy as np
import pytest

import clc_hip
from gpu_helpers import TYPE_OF, dev, host, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def scan_both(ctx, x: np.ndarray, op, identity, exclusive):
    import torch
    L = clc_hip.lib()
    t = TYPE_OF[x.dtype]
    xd = dev(x)
    ref = torch.empty_like(xd)
    got = torch.empty_like(xd)
    torch.cuda.synchronize()
    clc_hip.check(L.clc_scan(ctx.handle, ctypes.c_void_p(xd.data_ptr()),
                             ctypes.c_void_p(ref.data_ptr()), x.size, t,
                             op.encode(), identity.encode(), int(exclusive)))
    torch.cuda.synchronize()
    before = ctx.memory_stats()
    clc_hip.check(L.clc_scan_noalloc(ctx.handle, ctypes.c_void_p(xd.data_ptr()),
                                     ctypes.c_void_p(got.data_ptr()), x.size, t,
                                     op.encode(), identity.encode(), int(exclusive)))
    torch.cuda.synchronize()
    after = ctx.memory_stats()
    assert (after.total_allocations, after.reused, after.current_bytes) == \
        (before.total_allocations, before.reused, before.current_bytes)
    return host(ref, x.dtype), host(got, x.dtype)


CASES = [
    (np.uint32, "(a + b)", "0", True),
    (np.int64, "((a < b) ? b : a)", "LONG_MIN", False),
    (np.uint8, "(a ^ b)", "0", False),
    (np.float32, "(a + b)", "0", False),
    (np.uint16, "(a * b)", "1", True),
    (np.float64, "((b < a) ? b : a)", "INFINITY", True),
]


@pytest.mark.parametrize("dtype,op,identity,exclusive", CASES)
def test_noalloc_scan_matches_scan_twice(ctx, dtype, op, identity, exclusive):
    rng = np.random.default_rng(5)
    for n in (3_000_017, 70_001):  # the second call is smaller
        say("scan_noalloc", np.dtype(dtype).name, op, exclusive, n)
        if np.dtype(dtype).kind == "f":
            x = rng.integers(-1000, 1000, n).astype(dtype)  # exact sums
        elif dtype == np.uint16:
            x = rng.integers(0, 4, n).astype(dtype) * 2 + 1  # odd factors
        else:
            info = np.iinfo(dtype)
            x = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
        ref, got = scan_both(ctx, x, op, identity, exclusive)
        np.testing.assert_array_equal(got.view(np.uint8), ref.view(np.uint8))


def test_noalloc_scan_in_place_and_edges(ctx):
    import torch
    L = clc_hip.lib()
    say("scan_noalloc in place, n == 0, refusals")
    x = np.arange(100_000, dtype=np.uint32) % 251
    xd = dev(x)
    torch.cuda.synchronize()
    clc_hip.check(L.clc_scan_noalloc(ctx.handle, ctypes.c_void_p(xd.data_ptr()),
                                     ctypes.c_void_p(xd.data_ptr()), x.size, clc_hip.U32,
                                     b"(a + b)", b"0", 0))
    np.testing.assert_array_equal(host(xd, np.uint32), np.cumsum(x, dtype=np.uint32))
    assert L.clc_scan_noalloc(ctx.handle, None, None, 0, clc_hip.U32, b"(a + b)", b"0",
                              0) == clc_hip.SUCCESS
    assert L.clc_scan_noalloc(ctx.handle, None, None, 10, clc_hip.U32, b"(a + b)", b"0",
                              0) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_scan_noalloc(ctx.handle, ctypes.c_void_p(xd.data_ptr()),
                              ctypes.c_void_p(xd.data_ptr()), 10, clc_hip.U32,
                              b"a + b + 1", b"0", 0) == clc_hip.ERROR_UNSUPPORTED
    assert L.clc_scan_noalloc(ctx.handle, ctypes.c_void_p(xd.data_ptr()),
                              ctypes.c_void_p(xd.data_ptr()), 1 << 31, clc_hip.U32,
                              b"(a + b)", b"0", 0) == clc_hip.ERROR_INVALID_ARGUMENT
