"""Sizes around and above 2^30 keys on the GPU, through the C ABI.

clc_sort takes up to 2^31-1 keys per call (the reference's check_size,
lib/src/Compute/Detail.hpp:18-23). From 2^30 keys on, a digit count no
longer fits the 30-bit count of a 32-bit look-back status word, so the sort
switches to 64-bit status words; the segmented sort keeps 32-bit words and
leaves a segment of 2^30 keys or more untouched, as its contract says; the
learned sort hands such inputs to the plain sort. Results are checked by the
size-independent stable-sort properties (keys ascending, values a
permutation, keys[i] == original[values[i]], ties ordered by value), which
determine a stable sort's output uniquely, and the untouched segment by
equality with a copy taken before the call."""
import gc

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import dev, host, say
from test_sort_gpu import check_stable_sort_properties

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


@pytest.fixture(autouse=True)
def free_memory():
    yield
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_sort_u32_pairs_2p30_plus_ragged_tail(ctx):
    import torch
    n = (1 << 30) + 4099
    say("sort u32 pairs, 64-bit status words", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(30)
    keys = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32,
                         device="cuda", generator=g)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    ctx.sort(keys, vals, key_type=clc_hip.U32)
    torch.cuda.synchronize()
    check_stable_sort_properties(orig, keys, vals)


@pytest.mark.parametrize("with_values", [True, False])
def test_sort_u32_2p30_minus_one(ctx, with_values):
    """The largest size on 32-bit status words: the write-out's 32-bit byte
    offsets from the output bases reach 2^32 - 4 here."""
    import torch
    n = (1 << 30) - 1
    say("sort u32, 32-bit status words, values" if with_values else "sort u32 keys only", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(34)
    keys = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32,
                         device="cuda", generator=g)
    orig = keys.clone()
    if with_values:
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        ctx.sort(keys, vals, key_type=clc_hip.U32)
        torch.cuda.synchronize()
        check_stable_sort_properties(orig, keys, vals)
    else:
        ctx.sort(keys, key_type=clc_hip.U32)
        torch.cuda.synchronize()
        # Keys only: ascending, and the same multiset as the input.
        ku = keys.to(torch.int64) & 0xFFFFFFFF
        assert bool((ku[1:] >= ku[:-1]).all())
        ou = orig.to(torch.int64) & 0xFFFFFFFF
        assert bool(torch.equal(torch.sort(ou).values, ku))


def test_sort_u64_pairs_2p30_plus_one(ctx):
    import torch
    n = (1 << 30) + 1
    say("sort u64 pairs in [0,2^40), 64-bit status words", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    keys = torch.randint(0, 1 << 40, (n,), dtype=torch.int64, device="cuda", generator=g)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    ctx.sort(keys, vals, key_type=clc_hip.U64)
    torch.cuda.synchronize()
    check_stable_sort_properties(orig, keys, vals)


def test_learned_sort_at_2p30_runs_the_plain_sort(ctx):
    import torch
    n = (1 << 30) + 1
    say("learned sort u64 pairs at 2^30+1 (plain sort)", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(32)
    keys = torch.randint(0, 1 << 48, (n,), dtype=torch.int64, device="cuda", generator=g)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    st = ctx.learned_sort(keys, vals, key_type=clc_hip.U64)
    torch.cuda.synchronize()
    assert st["used_model"] == 0
    check_stable_sort_properties(orig, keys, vals)


def test_segment_of_2p30_keys_is_left_untouched(ctx):
    import torch
    big = (1 << 30) + 1
    sizes = [1000, big, 1999]
    n = sum(sizes)
    say("segmented sort with a segment of 2^30+1 keys", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(33)
    keys = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32,
                         device="cuda", generator=g)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    orig = keys.clone()
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    ctx.segmented_sort(keys, vals, dev(off), key_type=clc_hip.U32)
    torch.cuda.synchronize()
    b, e = int(off[1]), int(off[2])
    assert bool(torch.equal(keys[b:e], orig[b:e]))
    assert bool(torch.equal(vals[b:e], torch.arange(b, e, dtype=torch.int32, device="cuda")))
    # The small segments either side are sorted as the oracle sorts them.
    for lo, hi in ((0, b), (e, n)):
        k0 = host(orig[lo:hi].contiguous(), np.uint32)
        ek, ev = oracle.radix_sort(k0, np.arange(lo, hi, dtype=np.uint32))
        np.testing.assert_array_equal(host(keys[lo:hi].contiguous(), np.uint32), ek)
        np.testing.assert_array_equal(host(vals[lo:hi].contiguous(), np.uint32), ev)


def test_segment_of_2p30_minus_one_keys_is_sorted(ctx):
    """The largest segment the segmented sort takes: its large path's 32-bit
    byte offsets from the segment's base reach 2^32 - 4."""
    import torch
    big = (1 << 30) - 1
    sizes = [1000, big, 1999]
    n = sum(sizes)
    say("segmented sort with a segment of 2^30-1 keys", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(35)
    keys = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32,
                         device="cuda", generator=g)
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    orig = keys.clone()
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    ctx.segmented_sort(keys, vals, dev(off), key_type=clc_hip.U32)
    torch.cuda.synchronize()
    b, e = int(off[1]), int(off[2])
    check_stable_sort_properties(orig[b:e], keys[b:e], vals[b:e] - b)
    for lo, hi in ((0, b), (e, n)):
        k0 = host(orig[lo:hi].contiguous(), np.uint32)
        ek, ev = oracle.radix_sort(k0, np.arange(lo, hi, dtype=np.uint32))
        np.testing.assert_array_equal(host(keys[lo:hi].contiguous(), np.uint32), ek)
        np.testing.assert_array_equal(host(vals[lo:hi].contiguous(), np.uint32), ev)
