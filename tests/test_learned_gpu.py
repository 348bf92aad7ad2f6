"""Learned-bucket sort on the GPU through the C ABI (clc_learned_sort).

Every output must be bit-identical to the oracle's stable sort (keys and
value permutation) and, at full size, to the library's plain radix sort
(clc_sort), the bit-exact reference. Covers the three skewed u64 key
distributions of the learned configuration (Zipf s=1.2 over [1, 2^40],
lognormal(20, 2), a mixture of 8 Gaussians), uniform keys, signed and float
keys in both directions, keys-only, degenerate inputs (one repeated key,
two values), the small-input path and the model statistics."""
import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say
from test_sort_gpu import check_stable_sort_properties

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = clc_hip.Context()
    yield c
    c.close()


def learned(ctx, keys, vals=None, descending=False):
    kd = dev(keys)
    vd = None if vals is None else dev(vals)
    st = ctx.learned_sort(kd, vd, descending, key_type=TYPE_OF[keys.dtype])
    return host(kd, keys.dtype), (None if vd is None else host(vd, np.uint32)), st


def check(ctx, keys, with_values=True, descending=False):
    vals = np.arange(keys.size, dtype=np.uint32) if with_values else None
    k, v, st = learned(ctx, keys, vals, descending)
    ek, ev = oracle.radix_sort(keys, vals, descending=descending, threads=8)
    np.testing.assert_array_equal(k.view(np.uint8), ek.view(np.uint8))
    if with_values:
        np.testing.assert_array_equal(v, ev)
    return st


DISTS = {
    "zipf": lambda n, seed: oracle.zipf_keys(n, 1.2, 1 << 40, seed),
    "lognormal": lambda n, seed: oracle.lognormal_keys(n, 20.0, 2.0, seed),
    "gmm": lambda n, seed: oracle.gaussian_mixture_keys(n, 8, 2.0**60, 2.0**40, seed),
    "uniform48": lambda n, seed: oracle.pcg32_u64(seed, 3, n, bits=48),
}


@pytest.mark.parametrize("dist", list(DISTS))
def test_distributions_u64_pairs(ctx, dist):
    n = 1 << 22
    say("learned", dist, n)
    st = check(ctx, DISTS[dist](n, 5))
    say("stats", st)
    # The model is used (else this would compare the plain sort with the
    # oracle again), with one bucket per ~4096 keys.
    assert st["used_model"] == 1
    assert st["num_buckets"] == 1024
    # Every 100th key up to 65,536 samples, every 400th beyond once larger.
    assert st["sample_size"] == max(n // 400, min(n // 100, 65536))
    assert 0.0 <= st["fallback_fraction"] <= 1.0


def test_gmm_uses_the_model(ctx):
    n = 1 << 22
    say("learned gmm model use")
    st = check(ctx, DISTS["gmm"](n, 9))
    assert st["used_model"] == 1
    assert st["num_buckets"] >= 512
    assert st["max_bucket_overfill"] < 16


@pytest.mark.parametrize("dtype,descending", [(np.uint32, False), (np.int32, True),
                                              (np.float32, False), (np.float32, True),
                                              (np.int64, False), (np.float64, True)])
def test_other_key_types(ctx, dtype, descending):
    n = 3 << 20
    say("learned", np.dtype(dtype).name, descending)
    rng = np.random.default_rng(7)
    if np.dtype(dtype).kind == "f":
        keys = (rng.standard_normal(n) * 1000).astype(dtype)
        keys[::97] = 0.0
        keys[::89] = -0.0
    else:
        info = np.iinfo(dtype)
        keys = rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True)
    check(ctx, keys, True, descending)


def test_keys_only(ctx):
    say("learned keys only")
    check(ctx, DISTS["lognormal"](1 << 21, 2), with_values=False)


@pytest.mark.parametrize("case", ["all_equal", "two_values", "sorted", "small"])
def test_degenerate_inputs(ctx, case):
    say("learned degenerate", case)
    n = 1000 if case == "small" else 1 << 20
    base = DISTS["gmm"](n, 4)
    keys = {"all_equal": np.full(n, 12345, np.uint64),
            "two_values": np.where(base & np.uint64(1), 7, 1 << 50).astype(np.uint64),
            "sorted": np.sort(base), "small": base}[case]
    check(ctx, keys)


@pytest.mark.parametrize("dist", ["zipf_s1.2_2^40", "lognormal_20_2", "gauss_mix_8"])
def test_full_size_2p28(ctx, dist):
    """The learned configuration at full size: 2^28 u64 keys with u32 index
    values from each skewed distribution, generated on the GPU as the
    benchmark generates them. The output must satisfy the stable-sort
    properties against the input (which fix it uniquely) and equal the plain
    radix sort's bit for bit; the model must be used, with 65,536 buckets
    (the two-byte bucket id, two id passes)."""
    import torch
    import bench
    n = 1 << 28
    say("learned full size 2^28 u64 pairs", dist)
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 + list(bench.LEARNED_DISTS).index(dist))
    keys = bench.LEARNED_DISTS[dist](torch, n, g)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    k2, v2 = keys.clone(), vals.clone()
    st = ctx.learned_sort(keys, vals, key_type=clc_hip.U64)
    ctx.sort(k2, v2, key_type=clc_hip.U64)
    torch.cuda.synchronize()
    say("stats", st)
    assert st["used_model"] == 1
    assert st["num_buckets"] == 65536
    same_keys = bool(torch.equal(keys, k2))
    same_vals = bool(torch.equal(vals, v2))
    assert same_keys and same_vals
    del k2, v2
    check_stable_sort_properties(orig, keys, vals)
