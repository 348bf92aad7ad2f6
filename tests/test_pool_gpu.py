"""The per-GPU scratch pool on the device, through the C ABI.

Every context on a GPU shares that GPU's pool, and a block parked by one
context's stream may be taken by another's: the pool then orders the taker
after the owner on the device. These tests sort on two contexts bound to two
torch streams back to back, so the second sort reuses the first one's
scratch while the first may still be running, and check both results
against the oracle; they also check the pool's counters and that buffers
and graphs may hand their blocks back after a context is gone."""
import ctypes

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import dev, host, say

pytestmark = pytest.mark.gpu


def test_two_contexts_share_scratch_in_stream_order():
    import torch
    say("pool: two contexts on two streams reuse one scratch block")
    n = 1 << 22
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    c1, c2 = clc_hip.Context(stream=s1), clc_hip.Context(stream=s2)
    try:
        ka = oracle.pcg32_u32(42, 54, n)
        kb = oracle.pcg32_u32(7, 11, n)
        vals = np.arange(n, dtype=np.uint32)
        a_k, a_v, b_k, b_v = dev(ka), dev(vals), dev(kb), dev(vals)
        torch.cuda.synchronize()
        # Start from an empty pool, so each sort's scratch is the one block
        # the other stream parked last.
        clc_hip.check(clc_hip.lib().clc_context_release_cache(c1.handle))
        before = c1.memory_stats()
        for rep in range(3):
            if rep > 0:
                with torch.cuda.stream(s1):
                    a_k.copy_(dev(ka))
                    a_v.copy_(dev(vals))
                with torch.cuda.stream(s2):
                    b_k.copy_(dev(kb))
                    b_v.copy_(dev(vals))
                torch.cuda.synchronize()
            c1.sort(a_k, a_v, key_type=clc_hip.U32)  # queued on s1; its scratch is parked on s1
            c2.sort(b_k, b_v, key_type=clc_hip.U32)  # same size on s2: takes s1's block
            torch.cuda.synchronize()
            ek, ev = oracle.radix_sort(ka, vals, threads=8)
            np.testing.assert_array_equal(host(a_k, np.uint32), ek)
            np.testing.assert_array_equal(host(a_v, np.uint32), ev)
            ek, ev = oracle.radix_sort(kb, vals, threads=8)
            np.testing.assert_array_equal(host(b_k, np.uint32), ek)
            np.testing.assert_array_equal(host(b_v, np.uint32), ev)
        after = c2.memory_stats()
        # The two contexts report the same (device-wide) pool.
        same = c1.memory_stats()
        assert same.reused == after.reused and same.current_bytes == after.current_bytes
        assert after.cross_stream - before.cross_stream >= 5
        assert after.reused >= after.cross_stream >= after.stream_waits
        assert after.current_bytes >= after.cached_bytes
        assert after.peak_bytes >= after.current_bytes
    finally:
        c1.close()
        c2.close()


def test_buffer_outlives_its_context():
    import torch
    say("pool: buffer released after its context")
    lib = clc_hip.lib()
    c = clc_hip.Context(stream=torch.cuda.Stream())
    buf = ctypes.c_void_p()
    clc_hip.check(lib.clc_buffer_alloc(c.handle, 1 << 20, ctypes.byref(buf)))
    data = np.arange(1 << 18, dtype=np.uint32)
    clc_hip.check(lib.clc_buffer_write(c.handle, buf, 0, data.nbytes,
                                       data.ctypes.data_as(ctypes.c_void_p)))
    other = clc_hip.Context(stream=torch.cuda.Stream())
    try:
        c.close()  # the buffer's home stream goes away
        back = np.empty_like(data)
        clc_hip.check(lib.clc_buffer_read(other.handle, buf, 0, back.nbytes,
                                          back.ctypes.data_as(ctypes.c_void_p)))
        np.testing.assert_array_equal(back, data)
        parked = other.memory_stats().cached_bytes
        lib.clc_buffer_release(buf)  # parked with no owner stream
        assert other.memory_stats().cached_bytes == parked + (1 << 20)
        # Every stream that parked a block is retired or this context's own,
        # so no take needs a wait.
        waits = other.memory_stats().stream_waits
        keys = oracle.pcg32_u32(3, 4, 1 << 17)
        kd = dev(keys)
        torch.cuda.synchronize()
        other.sort(kd, key_type=clc_hip.U32)
        other.finish()
        np.testing.assert_array_equal(host(kd, np.uint32), np.sort(keys, kind="stable"))
        assert other.memory_stats().stream_waits == waits
    finally:
        other.close()


def test_release_cache_returns_parked_blocks():
    import torch
    say("pool: release cache")
    c = clc_hip.Context(stream=torch.cuda.Stream())
    try:
        keys = dev(oracle.pcg32_u32(5, 6, 1 << 20))
        torch.cuda.synchronize()
        c.sort(keys, key_type=clc_hip.U32)
        c.finish()
        s = c.memory_stats()
        assert s.cached_bytes > 0
        clc_hip.check(clc_hip.lib().clc_context_release_cache(c.handle))
        s2 = c.memory_stats()
        assert s2.cached_bytes == 0
        assert s2.current_bytes == s.current_bytes - s.cached_bytes
    finally:
        c.close()
