"""The learned sort's single fused scatter (clc_learned_sort_single_scatter)
on the GPU, through the C ABI.

Per workgroup chunk the bucket counts (inference fused in), one exclusive
scan over them bucket-major, then one stable scatter into at most 4096
buckets. Every output must be bit-identical to the oracle's stable sort of
the reference algorithm (Algorithm.cpp:453-527), and at the largest sizes to
the library's plain radix sort and to the two-pass learned form. Covers the
three skewed distributions, sizes that end mid-tile and mid-chunk, every
key type in both directions, keys-only, and one repeated key."""
import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say
from test_learned_gpu import DISTS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def single(ctx, keys, vals=None, descending=False):
    kd = dev(keys)
    vd = None if vals is None else dev(vals)
    st = ctx.learned_sort(kd, vd, descending, key_type=TYPE_OF[keys.dtype], single_scatter=True)
    return host(kd, keys.dtype), (None if vd is None else host(vd, np.uint32)), st


def check(ctx, keys, with_values=True, descending=False):
    vals = np.arange(keys.size, dtype=np.uint32) if with_values else None
    k, v, st = single(ctx, keys, vals, descending)
    ek, ev = oracle.radix_sort(keys, vals, descending=descending, threads=8)
    np.testing.assert_array_equal(k.view(np.uint8), ek.view(np.uint8))
    if with_values:
        np.testing.assert_array_equal(v, ev)
    return st


@pytest.mark.parametrize("dist", list(DISTS))
def test_distributions_u64_pairs(ctx, dist):
    # 2^22 + 4099 keys: the last tile of the last chunk is partial.
    n = (1 << 22) + 4099
    say("single scatter", dist, n)
    st = check(ctx, DISTS[dist](n, 7))
    say("stats", st)
    assert st["used_model"] == 1
    assert st["num_buckets"] == 2048  # about 4096 keys a bucket, at most 4096
    assert 0.0 <= st["fallback_fraction"] <= 1.0


@pytest.mark.parametrize("n", [1 << 16, (1 << 16) + 1, 100003, 6144 * 256 + 1])
def test_sizes(ctx, n):
    say("single scatter size", n)
    st = check(ctx, DISTS["lognormal"](n, n % 97))
    assert st["num_buckets"] == 512


@pytest.mark.parametrize("dtype,descending,with_values", [
    (np.uint32, False, True), (np.int32, True, True), (np.float32, False, False),
    (np.float32, True, True), (np.uint64, True, False), (np.int64, False, True),
    (np.float64, True, True)])
def test_key_types(ctx, dtype, descending, with_values):
    n = 3 << 19
    rng = np.random.default_rng(31)
    dt = np.dtype(dtype)
    say("single scatter", dt.name, descending, with_values)
    if dt.kind == "f":
        keys = rng.lognormal(3.0, 2.0, n).astype(dt)
        keys[::2] *= -1
        keys[::17] = 0.0
        keys[::19] = -0.0
    else:
        info = np.iinfo(dt)
        keys = rng.integers(info.min // 2, info.max // 2, size=n, dtype=dt)
        keys[::5] = keys[11]
    check(ctx, keys, with_values, descending)


def test_one_repeated_key(ctx):
    n = 1 << 18
    keys = np.full(n, 12345, dtype=np.uint64)
    say("single scatter one key")
    check(ctx, keys)


def test_matches_two_pass_form_and_plain_sort_2p24(ctx):
    """At 2^24 keys (4096 buckets of about 4096 keys): the single scatter,
    the two-pass learned form and the plain radix sort agree bit for bit."""
    import torch
    n = 1 << 24
    for dist in ("zipf", "lognormal", "gmm"):
        say("single scatter vs two-pass vs plain", dist, n)
        keys = torch.from_numpy(DISTS[dist](n, 3).view(np.int64)).cuda()
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        out = []
        for mode in ("single", "two", "plain"):
            k, v = keys.clone(), vals.clone()
            if mode == "plain":
                ctx.sort(k, v, key_type=clc_hip.U64)
            else:
                st = ctx.learned_sort(k, v, key_type=clc_hip.U64, single_scatter=mode == "single")
                assert st["used_model"] == 1
                assert st["num_buckets"] == 4096
            torch.cuda.synchronize()
            out.append((k, v))
        for k, v in out[1:]:
            assert torch.equal(out[0][0], k) and torch.equal(out[0][1], v)
