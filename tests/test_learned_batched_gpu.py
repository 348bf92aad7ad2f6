"""Batched learned sort on the GPU through the C ABI
(clc_learned_segmented_sort): one learned model per segment.

Every output must be bit-identical to the oracle's per-segment stable sort
(keys and value permutation) and, at a larger size, to the library's plain
segmented sort. Covers the float32 segments of the batched learned
configuration (log-uniform sizes, each segment uniform[0,1), normal(0,1),
exponential(1) or a bimodal normal mix, with +-0 and duplicates), other
32-bit key types in both directions, keys-only, segment sizes around the
LDS capacity, degenerate and malformed segments, and the statistics."""
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


def loguniform_offsets(nseg, lo_log2, hi_log2, seed):
    rng = np.random.default_rng(seed)
    sizes = np.rint(2.0 ** rng.uniform(lo_log2, hi_log2, nseg)).astype(np.uint64)
    off = np.zeros(nseg + 1, dtype=np.uint64)
    np.cumsum(sizes, out=off[1:])
    return off


def c5_keys(off, seed):
    """float32 keys: segment i from distribution i % 4 (uniform[0,1),
    normal(0,1), exponential(1), bimodal normal mix), with +0, -0 and
    repeated values injected; no NaN."""
    rng = np.random.default_rng(seed)
    n = int(off[-1])
    keys = np.empty(n, dtype=np.float32)
    for i in range(len(off) - 1):
        b, e = int(off[i]), int(off[i + 1])
        m = e - b
        kind = i % 4
        if kind == 0:
            x = rng.random(m)
        elif kind == 1:
            x = rng.standard_normal(m)
        elif kind == 2:
            x = rng.exponential(1.0, m)
        else:
            x = np.where(rng.random(m) < 0.5, rng.normal(-3.0, 0.5, m), rng.normal(4.0, 1.0, m))
        x = x.astype(np.float32)
        if m >= 8:
            x[:: 37] = 0.0
            x[5:: 41] = -0.0
            x[7:: 29] = x[3]
        keys[b:e] = x
    return keys


def run(ctx, keys, vals, off, descending=False):
    kd = dev(keys)
    vd = None if vals is None else dev(vals)
    od = dev(off)
    st = ctx.learned_segmented_sort(kd, vd, od, descending, key_type=TYPE_OF[keys.dtype])
    return host(kd, keys.dtype), (None if vd is None else host(vd, np.uint32)), st


def check(ctx, keys, vals, off, descending=False):
    k, v, st = run(ctx, keys, vals, off, descending)
    ek, ev = oracle.segmented_sort(keys, vals, off, descending, threads=8)
    np.testing.assert_array_equal(k.view(np.uint8), ek.view(np.uint8))
    if vals is not None:
        np.testing.assert_array_equal(v, ev)
    return st


def test_c5_float_segments(ctx):
    off = loguniform_offsets(40, 10, 19, seed=3)
    n = int(off[-1])
    say("learned batched c5-style", len(off) - 1, "segments", n, "pairs")
    keys = c5_keys(off, 4)
    st = check(ctx, keys, np.arange(n, dtype=np.uint32), off)
    say("stats", st)
    assert st["sample_size"] > 0
    assert st["buckets"] >= len(off) - 1
    assert 0.0 <= st["fallback_fraction"] <= 1.0


@pytest.mark.parametrize("dtype,descending", [(np.float32, True), (np.uint32, False),
                                              (np.int32, True), (np.int32, False)])
def test_types_directions(ctx, dtype, descending):
    say("learned batched", np.dtype(dtype).name, descending)
    off = loguniform_offsets(24, 12, 18, seed=7)
    n = int(off[-1])
    rng = np.random.default_rng(11)
    if np.dtype(dtype).kind == "f":
        keys = c5_keys(off, 12)
    else:
        info = np.iinfo(dtype)
        keys = rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True)
        keys[::7] = keys[2]
    check(ctx, keys, rng.integers(0, 2**32, size=n, dtype=np.uint32), off, descending)


def test_keys_only(ctx):
    say("learned batched keys only")
    off = loguniform_offsets(16, 14, 18, seed=9)
    check(ctx, c5_keys(off, 2), None, off)


def test_sizes_around_capacity(ctx):
    """Segments just below, at and above the LDS capacity, and one with a
    partial last onesweep tile of one key."""
    sizes = [0, 1, 2, 1000, 16383, 16384, 16385, 20000, 14336, 14337, 28673, 100000, 300001]
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    np.cumsum(np.array(sizes, dtype=np.uint64), out=off[1:])
    n = int(off[-1])
    say("learned batched capacity boundaries", n)
    check(ctx, c5_keys(off, 5), np.arange(n, dtype=np.uint32), off)


@pytest.mark.parametrize("case", ["all_equal", "two_values", "sorted", "reversed"])
def test_degenerate_segments(ctx, case):
    say("learned batched degenerate", case)
    off = loguniform_offsets(8, 15, 18, seed=13)
    n = int(off[-1])
    base = c5_keys(off, 6)
    if case == "all_equal":
        keys = np.full(n, 2.5, np.float32)
    elif case == "two_values":
        keys = np.where(np.arange(n) % 3 == 0, -1.0, 7.0).astype(np.float32)
    else:
        keys = base.copy()
        for i in range(len(off) - 1):
            b, e = int(off[i]), int(off[i + 1])
            seg = np.sort(keys[b:e])
            keys[b:e] = seg if case == "sorted" else seg[::-1]
    st = check(ctx, keys, np.arange(n, dtype=np.uint32), off)
    say("stats", st)


def test_malformed_segments_untouched(ctx):
    say("learned batched malformed offsets")
    n = 200000
    keys = c5_keys(np.array([0, n], dtype=np.uint64), 8)
    vals = np.arange(n, dtype=np.uint32)
    # [60000, 250000) and [250000, 260000) end past n: skipped, keys untouched
    off = np.array([0, 60000, 250000, 260000], dtype=np.uint64)
    k, v, _ = run(ctx, keys, vals, off)
    ek, ev = keys.copy(), vals.copy()
    order = np.argsort(oracle.transformed_keys(keys[:60000]), kind="stable")
    ek[:60000] = keys[:60000][order]
    ev[:60000] = vals[:60000][order]
    np.testing.assert_array_equal(k.view(np.uint8), ek.view(np.uint8))
    np.testing.assert_array_equal(v, ev)


def test_matches_plain_segmented_sort(ctx):
    """Larger batch: bit-identical to the plain segmented sort."""
    import torch
    off = loguniform_offsets(256, 10, 20, seed=21)
    n = int(off[-1])
    say("learned batched vs plain segmented", n)
    keys = dev(c5_keys(off, 22))
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    od = dev(off)
    k2, v2 = keys.clone(), vals.clone()
    st = ctx.learned_segmented_sort(keys, vals, od, key_type=clc_hip.F32)
    ctx.segmented_sort(k2, v2, od, key_type=clc_hip.F32)
    torch.cuda.synchronize()
    same = bool(torch.equal(keys, k2)) and bool(torch.equal(vals, v2))
    say("stats", st)
    assert same


def test_errors(ctx):
    import torch
    keys = torch.zeros(16, dtype=torch.int64, device="cuda")
    od = dev(np.array([0, 16], dtype=np.uint64))
    with pytest.raises(clc_hip.ClcError):
        ctx.learned_segmented_sort(keys, None, od, key_type=clc_hip.U64)


def check_segmented_properties(orig, off_d, keys, vals, chunk=1 << 27):
    """Size-independent check that (keys, vals) is the per-segment stable
    ascending sort of the float32 keys orig, with vals the global index of
    each key: within each segment the order images are non-decreasing;
    every value lies in its own element's segment; the values form a
    permutation; keys[i] == orig[vals[i]] bit for bit; and equal keys in a
    segment keep index order. These facts fix each segment's stable sort
    uniquely. Runs on the GPU in chunks to bound the temporaries."""
    import torch
    n = keys.numel()
    kb = keys.view(torch.int32)
    ob = orig.view(torch.int32)
    # order image of float32 bits (the reference's KEY_TRANSFORM), as int64
    def image(bits):
        b = bits.to(torch.int64) & 0xFFFFFFFF
        return torch.where(b >= 0x80000000, b ^ 0xFFFFFFFF, b | 0x80000000)
    seen = torch.zeros(n, dtype=torch.bool, device="cuda")
    ok = True
    for a in range(0, n, chunk):
        b = min(n, a + chunk + 1)  # one element of overlap for the pairwise tests
        idx = torch.arange(a, b, dtype=torch.int64, device="cuda")
        seg = torch.searchsorted(off_d, idx, right=True) - 1
        lo, hi = off_d[seg], off_d[seg + 1]
        v = vals[a:b].to(torch.int64) & 0xFFFFFFFF
        ok = ok and bool(((v >= lo) & (v < hi)).all())
        ok = ok and bool(torch.equal(ob[v], kb[a:b]))
        img = image(kb[a:b])
        same_seg = seg[1:] == seg[:-1]
        ok = ok and bool((img[1:] >= img[:-1])[same_seg].all())
        ties = same_seg & (img[1:] == img[:-1])
        ok = ok and bool((v[1:][ties] > v[:-1][ties]).all())
        seen[v] = True
        del idx, seg, lo, hi, v, img, same_seg, ties
        if not ok:
            break
    return ok and bool(seen.all())


def test_c5_full_configuration(ctx):
    """The batched learned configuration at full size, as the benchmark runs
    it: 4096 float32 segments with sizes 2^U(10,22) (about 2.08e9 pairs),
    generated on the GPU, u32 values = global index. Every segment is
    checked on the GPU by the stable-sort properties; the three largest
    segments are also compared with the oracle's per-segment sort."""
    import torch
    import bench
    nseg = 4096
    rng = np.random.default_rng(4096)
    sizes = np.rint(2.0 ** rng.uniform(10.0, 22.0, nseg)).astype(np.int64)
    off = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    total = int(off[-1])
    say("learned batched full configuration", nseg, "segments,", total, "pairs")
    g = torch.Generator(device="cuda")
    g.manual_seed(4096)
    keys = bench.c5_keys_gpu(torch, off, g)
    orig = keys.clone()
    vals = torch.arange(total, dtype=torch.int32, device="cuda")
    off_d = torch.from_numpy(off).cuda()
    st = ctx.learned_segmented_sort(keys, vals, off_d, key_type=clc_hip.F32)
    torch.cuda.synchronize()
    say("stats", st)
    assert st["buckets"] > nseg
    assert check_segmented_properties(orig, off_d, keys, vals)
    for i in np.argsort(sizes)[-3:]:
        b, e = int(off[i]), int(off[i + 1])
        say("oracle check of segment", int(i), "size", e - b)
        k0 = orig[b:e].cpu().numpy()
        ek, ev = oracle.radix_sort(k0, np.arange(b, e, dtype=np.uint32), threads=8)
        np.testing.assert_array_equal(keys[b:e].cpu().numpy().view(np.uint32),
                                      ek.view(np.uint32))
        np.testing.assert_array_equal(vals[b:e].cpu().numpy().view(np.uint32), ev)
