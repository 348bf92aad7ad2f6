"""Segmented sort parity on the GPU through the C ABI (clc_segmented_sort
and the hipGraph executor), against the oracle's per-segment restatement
of the reference sort. Bit-exact keys and value permutations.

Covers: every size class (LDS configurations, the 256 x 24 shape for
4097..6144 keys and the multi-workgroup onesweep path), class boundaries,
empty and single-key segments, malformed
segments left untouched, every key type, both directions, keys-only, the
lognormal batched configuration at full size (65,536 segments), tuned
tables, and graph replays with eager work in between."""
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


def run(ctx, keys, vals, off, descending=False):
    kd = dev(keys)
    vd = None if vals is None else dev(vals)
    od = dev(off.astype(np.uint64))
    ctx.segmented_sort(kd, vd, od, descending, key_type=TYPE_OF[keys.dtype])
    return host(kd, keys.dtype), (None if vd is None else host(vd, np.uint32))


def check(ctx, keys, vals, off, descending=False, threads=8):
    k, v = run(ctx, keys, vals, off, descending)
    ek, ev = oracle.segmented_sort(keys, vals, off, descending, threads=threads)
    np.testing.assert_array_equal(k.view(np.uint8), ek.view(np.uint8))
    if vals is not None:
        np.testing.assert_array_equal(v, ev)


def boundary_offsets():
    sizes = [0, 1, 2, 3, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025,
             2047, 2048, 2049, 4095, 4096, 4097, 5119, 5120, 5121, 6143, 6144, 6145,
             8191, 8192, 8193,
             16383, 16384, 16385,
             20000, 40961, 100000, 262144, 300001]
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    np.cumsum(np.array(sizes, dtype=np.uint64), out=off[1:])
    return off


def test_class_boundaries_u32_pairs(ctx):
    say("segmented class boundaries u32 pairs")
    off = boundary_offsets()
    n = int(off[-1])
    keys = oracle.pcg32_u32(3, 4, n)
    vals = np.arange(n, dtype=np.uint32)
    check(ctx, keys, vals, off)


@pytest.mark.parametrize("bits", [16, 48, 49, 64])
def test_class_boundaries_u64_pairs(ctx, bits):
    """64-bit keys with values are sorted in LDS as packed (digits << 16 |
    position) words, 48 bits of the key range at a time: ranges of up to
    48 bits and just past it take different paths."""
    say("segmented class boundaries u64 pairs", bits)
    off = boundary_offsets()
    n = int(off[-1])
    keys = oracle.pcg32_u64(5, bits, n, bits=bits)
    keys[::5] = keys[3]  # equal keys: stability
    check(ctx, keys, np.arange(n, dtype=np.uint32), off)


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.float32,
                                   np.uint64, np.int64, np.float64])
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("with_values", [False, True])
def test_types_directions(ctx, dtype, descending, with_values):
    say("segmented", np.dtype(dtype).name, descending, with_values)
    off = oracle.lognormal_segments(2000, np.log(600), 1.6, 0, 200000, seed=9)
    n = int(off[-1])
    rng = np.random.default_rng(2)
    dt = np.dtype(dtype)
    if dt.kind == "f":
        keys = (rng.standard_normal(n) * 100).astype(dt)
        keys[::13] = 0.0
        keys[::17] = -0.0
        keys[::19] = keys[5]
    else:
        info = np.iinfo(dt)
        keys = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
        keys[::3] = keys[7]
    vals = rng.integers(0, 2**32, size=n, dtype=np.uint32) if with_values else None
    check(ctx, keys, vals, off, descending)


def test_duplicates_and_sorted_segments(ctx):
    say("segmented duplicates / presorted")
    off = oracle.lognormal_segments(500, np.log(3000), 1.2, 1, 60000, seed=5)
    n = int(off[-1])
    keys = (oracle.pcg32_u32(1, 1, n) % 7).astype(np.uint32)
    keys[: n // 2] = np.sort(keys[: n // 2])
    check(ctx, keys, np.arange(n, dtype=np.uint32), off)


def test_malformed_segments_untouched(ctx):
    say("segmented malformed offsets")
    n = 10000
    keys = oracle.pcg32_u32(8, 8, n)
    vals = np.arange(n, dtype=np.uint32)
    # [0,3000) ok; [3000,3000) empty; [3000,9000) ok; [9000,12000) past the end
    off = np.array([0, 3000, 3000, 9000, 12000], dtype=np.uint64)
    k, v = run(ctx, keys, vals, off)
    ek, ev = oracle.radix_sort(keys[:3000], vals[:3000])
    np.testing.assert_array_equal(k[:3000], ek)
    np.testing.assert_array_equal(v[:3000], ev)
    ek2, ev2 = oracle.radix_sort(keys[3000:9000], vals[3000:9000])
    np.testing.assert_array_equal(k[3000:9000], ek2)
    np.testing.assert_array_equal(k[9000:], keys[9000:])
    np.testing.assert_array_equal(v[9000:], vals[9000:])
    # [0,5000) ok; [5000,4000) inverted: left alone
    off = np.array([0, 5000, 4000], dtype=np.uint64)
    k, v = run(ctx, keys, vals, off)
    np.testing.assert_array_equal(k[:5000], oracle.radix_sort(keys[:5000])[0])
    np.testing.assert_array_equal(k[5000:], keys[5000:])


def test_c3_full_size_lognormal_65536(ctx):
    """The batched configuration at full size, through both entry points:
    one direct clc_segmented_sort call and two replays of the captured
    hipGraph (with eager work between them), each against the oracle."""
    import torch
    say("segmented C3: 65536 lognormal segments (mu ln 512, sigma 1.5)")
    off = oracle.lognormal_segments(65536, np.log(512), 1.5, 1, 1 << 20, seed=3)
    n = int(off[-1])
    keys = oracle.pcg32_u32(42, 56, n)
    seg_of = np.repeat(np.arange(off.size - 1), np.diff(off).astype(np.int64))
    vals = (np.arange(n, dtype=np.uint64) - off[seg_of]).astype(np.uint32)
    del seg_of
    ek, ev = oracle.segmented_sort(keys, vals, off, threads=16)
    k, v = run(ctx, keys, vals, off)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
    say("segmented C3: the same sort as one hipGraph launch")
    kd, vd, od = dev(keys), dev(vals), dev(off)
    g = ctx.segmented_sort_graph(kd, vd, od, key_type=clc_hip.U32)
    try:
        for rep in range(2):
            if rep:
                kd.copy_(dev(keys))
                vd.copy_(dev(vals))
                ctx.sort(dev(keys[:5000]), None, key_type=clc_hip.U32)
            g.launch()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(host(kd, np.uint32), ek)
            np.testing.assert_array_equal(host(vd, np.uint32), ev)
    finally:
        g.close()


def test_tuned_table_and_rerun(ctx):
    say("segmented autotune u32 pairs")
    ctx.segmented_sort_tune(clc_hip.U32, True)
    t = ctx.segmented_sort_table(clc_hip.U32, True)
    assert len(t) == 32 and all(0 <= c <= 7 for c in t)  # 7: the onesweep path
    off = boundary_offsets()
    n = int(off[-1])
    keys = oracle.pcg32_u32(6, 6, n)
    check(ctx, keys, np.arange(n, dtype=np.uint32), off)


def test_graph_replays(ctx):
    import torch
    say("segmented graph replays")
    off = oracle.lognormal_segments(3000, np.log(512), 1.5, 1, 1 << 18, seed=11)
    n = int(off[-1])
    kd = torch.empty(n, dtype=torch.int32, device="cuda")
    vd = torch.empty(n, dtype=torch.int32, device="cuda")
    od = dev(off)
    g = ctx.segmented_sort_graph(kd, vd, od, key_type=clc_hip.U32)
    vals = np.arange(n, dtype=np.uint32)
    for rep in range(4):
        keys = oracle.pcg32_u32(100 + rep, 1, n)
        kd.copy_(dev(keys))
        vd.copy_(dev(vals))
        # eager work between replays on the same stream
        ctx.sort(dev(keys[:1000]), None, key_type=clc_hip.U32)
        g.launch()
        torch.cuda.synchronize()
        ek, ev = oracle.segmented_sort(keys, vals, off, threads=8)
        np.testing.assert_array_equal(host(kd, np.uint32), ek)
        np.testing.assert_array_equal(host(vd, np.uint32), ev)
    g.close()


def test_errors(ctx):
    say("segmented error paths")
    L = clc_hip.lib()
    assert L.clc_segmented_sort(ctx.handle, ctypes.c_void_p(16), clc_hip.U16, None, 10,
                                ctypes.c_void_p(16), 2, 0) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_segmented_sort(ctx.handle, None, clc_hip.U32, None, 10, None, 2, 0) \
        == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_segmented_sort(ctx.handle, None, clc_hip.U32, None, 0, None, 0, 0) \
        == clc_hip.SUCCESS
