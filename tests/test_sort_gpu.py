"""Sort parity on the GPU, through the C ABI (clc_sort).

Mirrors the reference's test_sort (lib/test/Compute/main.cpp:246-312) and
the C-ABI sort checks (capi.c:133-151), then goes further: every key type
(including the 64-bit extension), both directions, keys-only and key-value,
ragged sizes around tile boundaries, skewed and degenerate inputs, and the
full 2^28 pair configuration. Results must be bit-identical to the oracle's
stable LSD sort (keys and value permutation)."""
import ctypes

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 7, 255, 256, 1000, 4097]


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    c = clc_hip.Context()
    yield c
    c.close()


def gpu_sort(ctx, keys: np.ndarray, values=None, descending=False):
    kd = dev(keys)
    vd = None if values is None else dev(values.astype(np.uint32))
    ctx.sort(kd, vd, descending=descending, key_type=TYPE_OF[keys.dtype])
    k = host(kd, keys.dtype)
    v = None if vd is None else host(vd, np.uint32)
    return k, v


def assert_same(a: np.ndarray, b: np.ndarray):
    np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("n", SIZES)
def test_reference_sort_vectors(ctx, golden, n):
    say("sort golden", n)
    keys = golden[f"sort_keys_{n}"]
    k, v = gpu_sort(ctx, keys, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(k, golden[f"sort_expect_keys_{n}"])
    np.testing.assert_array_equal(v, golden[f"sort_expect_values_{n}"])


@pytest.mark.parametrize("n", [100003, 1 << 20])
def test_reference_sizes_large(ctx, n):
    say("sort large reference size", n)
    keys = oracle.pcg32_u32(n + 2, 54, n)
    vals = np.arange(n, dtype=np.uint32)
    k, v = gpu_sort(ctx, keys, vals)
    ek, ev = oracle.radix_sort(keys, vals, threads=8)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)


def test_c1_keys_only_2p20_pcg32(ctx):
    """The reference test's largest size (main.cpp:54) as the CPU-runnable
    configuration states it: 2^20 raw pcg32 outputs (srandom(42, 54)),
    keys only, ascending."""
    n = 1 << 20
    say("sort C1 keys-only 2^20 pcg32")
    keys = oracle.pcg32_u32(42, 54, n)
    k, _ = gpu_sort(ctx, keys)
    np.testing.assert_array_equal(k, oracle.radix_sort(keys, threads=8)[0])


def test_reference_stability_vector(ctx, golden):
    say("sort stability 50000 keys in [0,15]")
    keys = golden["stable_keys_u8"].astype(np.uint32)
    k, v = gpu_sort(ctx, keys, np.arange(keys.size, dtype=np.uint32))
    np.testing.assert_array_equal(v, golden["stable_expect_values"])


def test_reference_signed_vector(ctx, golden):
    say("sort signed extremes")
    k, _ = gpu_sort(ctx, golden["ints_in"])
    np.testing.assert_array_equal(k, golden["ints_expect"])


def test_reference_float_descending_vector(ctx, golden):
    say("sort float descending")
    f = golden["floats_in"]
    k, _ = gpu_sort(ctx, f, descending=True)
    assert np.array_equal(k, golden["floats_expect_desc"])
    ek, _ = oracle.radix_sort(f, descending=True)
    assert_same(k, ek)


def test_capi_vector_keys_only(ctx, golden):
    say("sort capi vector")
    x = golden["capi_in"]
    k, _ = gpu_sort(ctx, x)
    assert np.all(k[1:] >= k[:-1])
    np.testing.assert_array_equal(k, np.sort(x, kind="stable"))


def make_keys(dtype, n, seed):
    rng = np.random.default_rng(seed)
    dt = np.dtype(dtype)
    if dt.kind == "f":
        keys = (rng.standard_normal(n) * 1e3).astype(dt)
        keys[: n // 50] = 0.0
        keys[n // 50: n // 25] = -0.0
        keys[n // 25: n // 25 + 10] = np.inf
        keys[n // 25 + 10: n // 25 + 20] = -np.inf
        keys[n // 10: n // 5] = keys[n // 2]
    else:
        info = np.iinfo(dt)
        keys = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
        keys[: n // 10] = keys[n // 2]
        keys[n // 10: n // 10 + 10] = info.min
        keys[n // 10 + 10: n // 10 + 20] = info.max
    return keys


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.float32,
                                   np.uint64, np.int64, np.float64])
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("with_values", [False, True])
def test_types_directions_vs_oracle(ctx, dtype, descending, with_values):
    n = 300007
    say("sort", np.dtype(dtype).name, "desc" if descending else "asc",
        "kv" if with_values else "keys", n)
    keys = make_keys(dtype, n, 11)
    vals = np.random.default_rng(3).integers(0, 2**32, size=n, dtype=np.uint32) \
        if with_values else None
    k, v = gpu_sort(ctx, keys, vals, descending)
    ek, ev = oracle.radix_sort(keys, vals, descending=descending, threads=8)
    assert_same(k, ek)
    if with_values:
        np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 255, 257, 4095, 4096, 4097,
                               8191, 8192, 8193, 12289, 65535, 65537])
def test_ragged_sizes(ctx, n):
    say("sort ragged", n)
    keys = oracle.pcg32_u32(n, 7, n)
    vals = np.arange(n, dtype=np.uint32)
    k, v = gpu_sort(ctx, keys, vals)
    ek, ev = oracle.radix_sort(keys, vals)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("case", ["all_equal", "two_values", "sorted", "reversed",
                                  "top_bits_only", "low_bits_only", "max_keys",
                                  "mixed_waves", "four_digits", "skew_after_first_round"])
def test_skewed_inputs(ctx, case):
    """Skewed keys, including inputs whose waves take both ranking paths
    (a wave whose first round has 16 or more lanes on one digit ranks by
    the ballot match, the others by LDS atomics): alternating runs of one
    wave's keys (64 x 26 at the default tile) with a constant or a random
    low byte, four low-byte values (16 lanes each, the threshold), and a
    uniform first round followed by constant rounds."""
    n = 200003
    say("sort skewed", case, n)
    base = oracle.pcg32_u32(5, 6, n)
    i = np.arange(n, dtype=np.uint32)
    wave_keys = 64 * 26
    keys = {
        "mixed_waves": np.where((i // wave_keys) % 2 == 0,
                                (base & np.uint32(0xFFFFFF00)) | np.uint32(5), base),
        "four_digits": (base & np.uint32(0xFFFFFF00)) | ((base >> 3) & np.uint32(3)),
        "skew_after_first_round": np.where(i % wave_keys < 64, base,
                                           base & np.uint32(0xFFFF0000)),
        "all_equal": np.full(n, 0xDEADBEEF, np.uint32),
        "two_values": np.where(base & 1, 3, 0xFFFFFFF0).astype(np.uint32),
        "sorted": np.sort(base),
        "reversed": np.sort(base)[::-1].copy(),
        "top_bits_only": base & np.uint32(0xFF000000),
        "low_bits_only": base & np.uint32(0xFF),
        "max_keys": np.full(n, 0xFFFFFFFF, np.uint32),
    }[case].astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = gpu_sort(ctx, keys, vals)
    ek, ev = oracle.radix_sort(keys, vals)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("dtype,bits", [(np.uint32, 8), (np.uint32, 16), (np.uint32, 24),
                                        (np.int32, 20), (np.uint64, 40), (np.uint64, 48),
                                        (np.uint64, 56), (np.int64, 12)])
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("with_values", [True, False])
def test_trivial_digit_places_skipped(ctx, dtype, bits, descending, with_values):
    """Keys confined to their low `bits` leave the upper digit places
    constant, so those passes are skipped on the device; odd and even counts
    of executed passes must both land the result in the caller's buffers
    (one executed pass through a final copy, three or more through the
    second scratch buffer)."""
    n = 150001
    say("sort trivial places", np.dtype(dtype).name, bits, descending, with_values)
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(dtype)
    vals = np.arange(n, dtype=np.uint32)
    k, v = gpu_sort(ctx, keys, vals if with_values else None, descending)
    ek, ev = oracle.radix_sort(keys, vals, descending=descending, threads=8)
    assert_same(k, ek)
    if with_values:
        np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("shape", range(11))
@pytest.mark.parametrize("dtype,bits", [(np.uint32, 24), (np.uint64, 40)])
@pytest.mark.parametrize("with_values", [True, False])
def test_every_tile_shape_odd_passes(ctx, monkeypatch, shape, dtype, bits, with_values):
    """Three and five executed digit places (an odd count, routed through the
    second scratch buffer) on every tile shape, the pipelined ones included."""
    monkeypatch.setenv("CLC_ONESWEEP_CFG", str(shape))
    for n in (1000, 300007):
        say("sort shape, odd passes", shape, np.dtype(dtype).name, bits, with_values, n)
        rng = np.random.default_rng(shape * 7 + n)
        keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(dtype)
        keys[:5] = (1 << bits) - 1  # the top executed place varies
        vals = np.arange(n, dtype=np.uint32)
        k, v = gpu_sort(ctx, keys, vals if with_values else None)
        ek, ev = oracle.radix_sort(keys, vals, threads=8)
        assert_same(k, ek)
        if with_values:
            np.testing.assert_array_equal(v, ev)


def test_unaligned_pointer(ctx):
    import torch
    n = 100001
    say("sort unaligned", n)
    keys = oracle.pcg32_u32(9, 9, n + 1)
    kd = dev(keys)
    vd = dev(np.arange(n + 1, dtype=np.uint32))
    sub_k, sub_v = kd[1:], vd[1:]
    clc_hip.check(clc_hip.lib().clc_sort(ctx.handle, ctypes.c_void_p(sub_k.data_ptr()),
                                         clc_hip.U32, ctypes.c_void_p(sub_v.data_ptr()),
                                         n, 0))
    torch.cuda.synchronize()
    ek, ev = oracle.radix_sort(keys[1:], np.arange(1, n + 1, dtype=np.uint32))
    np.testing.assert_array_equal(host(kd, np.uint32)[1:], ek)
    np.testing.assert_array_equal(host(vd, np.uint32)[1:], ev)
    assert host(kd, np.uint32)[0] == keys[0]


@pytest.mark.parametrize("shape", range(11))
@pytest.mark.parametrize("dtype,descending", [(np.uint32, False), (np.float32, True),
                                              (np.uint64, False), (np.int64, True)])
def test_every_tile_shape(ctx, monkeypatch, shape, dtype, descending):
    """Every onesweep tile shape the tuner can pick gives the oracle's output."""
    monkeypatch.setenv("CLC_ONESWEEP_CFG", str(shape))
    for n in (1000, 24577, 300007):
        say("sort shape", shape, np.dtype(dtype).name, descending, n)
        keys = make_keys(dtype, n, shape + n)
        vals = np.arange(n, dtype=np.uint32)
        k, v = gpu_sort(ctx, keys, vals, descending)
        ek, ev = oracle.radix_sort(keys, vals, descending=descending, threads=8)
        assert_same(k, ek)
        np.testing.assert_array_equal(v, ev)


def test_sort_tuner_picks_a_valid_shape(ctx):
    say("sort tuner u32 pairs")
    shape = ctx.sort_tune(clc_hip.U32, True)
    assert 0 <= shape < 11
    n = 100003
    keys = oracle.pcg32_u32(77, 1, n)
    vals = np.arange(n, dtype=np.uint32)
    k, v = gpu_sort(ctx, keys, vals)
    ek, ev = oracle.radix_sort(keys, vals)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)


def test_errors(ctx):
    L = clc_hip.lib()
    say("sort error paths")
    # 64-bit keys are an extension here; 16-bit keys stay rejected.
    assert L.clc_sort(ctx.handle, ctypes.c_void_p(16), clc_hip.U16, None, 10, 0) \
        == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_sort(ctx.handle, ctypes.c_void_p(16), clc_hip.U32, None,
                      1 << 31, 0) == clc_hip.ERROR_INVALID_ARGUMENT
    assert L.clc_sort(ctx.handle, None, clc_hip.U32, None, 0, 0) == clc_hip.SUCCESS


def check_stable_sort_properties(orig_keys, keys, vals):
    """Size-independent check that (keys, vals) is the stable ascending sort
    of orig_keys with vals = 0..n-1: keys ascending, vals a permutation,
    keys[i] == orig_keys[vals[i]], and ties ordered by vals. These four
    facts determine the stable sort uniquely."""
    import torch
    n = keys.numel()
    k = keys.to(torch.int64) & 0xFFFFFFFF if keys.dtype == torch.int32 else keys
    if keys.dtype == torch.int64:
        # unsigned 64-bit order via flipping the sign bit
        k = keys ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=keys.device)
    # Predicates are reduced to Python bools before assert, so a failure
    # never makes pytest render 2^28-element tensors.
    ascending = bool((k[1:] >= k[:-1]).all())
    v = vals.to(torch.int64) & 0xFFFFFFFF
    seen = torch.zeros(n, dtype=torch.bool, device=keys.device)
    seen[v] = True
    permutation = bool(seen.all())
    consistent = bool((orig_keys[v] == keys).all())
    ties = k[1:] == k[:-1]
    stable = bool((v[1:][ties] > v[:-1][ties]).all())
    assert ascending and permutation and consistent and stable, \
        (ascending, permutation, consistent, stable)


def test_full_size_2p28_pairs(ctx):
    import torch
    n = 1 << 28
    say("sort 2^28 u32 pairs (full configuration)")
    g = torch.Generator(device="cuda")
    g.manual_seed(42)
    keys = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32,
                         device="cuda", generator=g)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    ctx.sort(keys, vals, key_type=clc_hip.U32)
    torch.cuda.synchronize()
    check_stable_sort_properties(orig, keys, vals)


def test_full_size_2p27_u64_pairs(ctx):
    import torch
    n = 1 << 27
    say("sort 2^27 u64 keys in [0,2^48) with u32 values")
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    keys = torch.randint(0, 1 << 48, (n,), dtype=torch.int64, device="cuda",
                         generator=g)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    ctx.sort(keys, vals, key_type=clc_hip.U64)
    torch.cuda.synchronize()
    check_stable_sort_properties(orig, keys, vals)


def test_u64_vs_oracle_2p22(ctx):
    n = 1 << 22
    say("sort u64 [0,2^48) vs oracle", n)
    keys = oracle.pcg32_u64(42, 54, n, bits=48)
    vals = np.arange(n, dtype=np.uint32)
    k, v = gpu_sort(ctx, keys, vals)
    ek, ev = oracle.radix_sort(keys, vals, threads=8)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
