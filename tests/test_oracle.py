"""Pins the CPU oracle before anything is checked against it.

Each oracle function is checked against something that does not restate
its own formula: the published pcg32 demo sequence, the reference tests'
own inputs and std::stable_sort / std::sort results (tests/golden), numpy's
stable argsort, closed forms (numpy cumsum/sum with uint32 wrap), brute
force on small inputs, and planted mistakes that the checks must catch.
"""
import numpy as np
import pytest

import oracle

SIZES = [0, 1, 7, 255, 256, 1000, 4097]


# ---- pcg32 -------------------------------------------------------------------

def test_pcg32_published_demo_sequence():
    # pcg32-demo (O'Neill, pcg-c-basic) prints these for srandom(42, 54).
    got = oracle.pcg32_u32(42, 54, 6)
    assert [int(x) for x in got] == [0xa15c02b7, 0x7b47f409, 0xba1d3330,
                                     0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_pcg32_advance_matches_sequential():
    full = oracle.pcg32_u32(7, 11, 5000)
    for skip in (0, 1, 63, 1000, 4999):
        part = oracle.pcg32_u32(7, 11, 5000 - skip, skip=skip)
        np.testing.assert_array_equal(part, full[skip:])


def test_ints_range_bounds_and_byte_mask():
    v = oracle.ints_range(42, 54, 20000, 0, 255)
    assert v.min() >= 0 and v.max() <= 255
    # diff = 255 needs 8 bits and never rejects: the low byte of each draw.
    raw = oracle.pcg32_u32(42, 54, 20000)
    np.testing.assert_array_equal(v, (raw & 0xFF).astype(np.int32))
    w = oracle.ints_range(1, 2, 20000, -5, 9)
    assert w.min() == -5 and w.max() == 9


# ---- radix sort --------------------------------------------------------------

@pytest.mark.parametrize("n", SIZES)
def test_sort_reference_test_vectors(golden, n):
    keys = golden[f"sort_keys_{n}"]
    idx = np.arange(n, dtype=np.uint32)
    k, v = oracle.radix_sort(keys, idx)
    np.testing.assert_array_equal(k, golden[f"sort_expect_keys_{n}"])
    np.testing.assert_array_equal(v, golden[f"sort_expect_values_{n}"])


def test_sort_stability_reference_vector(golden):
    keys = golden["stable_keys_u8"].astype(np.uint32)
    k, v = oracle.radix_sort(keys, np.arange(keys.size, dtype=np.uint32))
    np.testing.assert_array_equal(v, golden["stable_expect_values"])
    np.testing.assert_array_equal(k, keys[golden["stable_expect_values"]])


def test_sort_signed_reference_vector(golden):
    k, _ = oracle.radix_sort(golden["ints_in"])
    np.testing.assert_array_equal(k, golden["ints_expect"])


def test_sort_float_descending_reference_vector(golden):
    f = golden["floats_in"]
    k, _ = oracle.radix_sort(f, descending=True)
    # The reference compares with ==, so -0 and +0 are alike (main.cpp:309).
    assert np.array_equal(k, golden["floats_expect_desc"])


@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.float32,
                                   np.uint64, np.int64, np.float64])
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("threads", [1, 3])
def test_sort_matches_numpy_stable_argsort(dtype, descending, threads):
    rng = np.random.default_rng(1234)
    n = 20011
    if np.dtype(dtype).kind == "f":
        keys = rng.standard_normal(n).astype(dtype) * 1e3
        keys[:50] = 0.0
        keys[50:100] = -0.0
        keys[100:110] = np.inf
        keys[110:120] = -np.inf
        keys[200:2200] = keys[3000]  # duplicates
    else:
        info = np.iinfo(dtype)
        keys = rng.integers(info.min, info.max, size=n, dtype=dtype, endpoint=True)
        keys[:300] = keys[400]
        keys[300:310] = info.min
        keys[310:320] = info.max
    vals = rng.integers(0, 2**32, size=n, dtype=np.uint32)
    perm = np.argsort(oracle.transformed_keys(keys, descending), kind="stable")
    k, v = oracle.radix_sort(keys, vals, descending=descending, threads=threads)
    np.testing.assert_array_equal(k.view(np.uint8), keys[perm].view(np.uint8))
    np.testing.assert_array_equal(v, vals[perm])


def test_sort_negative_zero_before_positive_zero():
    keys = np.array([0.0, -0.0, 1.0, -0.0, 0.0, -1.0], dtype=np.float32)
    k, v = oracle.radix_sort(keys, np.arange(6, dtype=np.uint32))
    assert v.tolist() == [5, 1, 3, 0, 4, 2]
    k, v = oracle.radix_sort(keys, np.arange(6, dtype=np.uint32), descending=True)
    assert v.tolist() == [2, 0, 4, 1, 3, 5]


def test_sort_small_brute_force():
    rng = np.random.default_rng(5)
    for n in range(0, 40):
        keys = rng.integers(0, 4, size=n, dtype=np.uint32) << np.uint32(29)
        vals = np.arange(n, dtype=np.uint32)
        pairs = sorted(zip(keys.tolist(), vals.tolist()), key=lambda p: p[0])
        k, v = oracle.radix_sort(keys, vals)
        assert list(zip(k.tolist(), v.tolist())) == pairs


def test_transform_mistakes_are_caught():
    """A planted sign mistake in the float transform must break the
    numpy comparison used above."""
    keys = np.array([-2.0, -1.0, 0.5, 3.0], dtype=np.float32)
    good = np.argsort(oracle.transformed_keys(keys), kind="stable")
    u = keys.view(np.uint32)
    bad_t = np.where((u & 0x80000000) != 0, u | 0x80000000, u)  # no inversion
    bad = np.argsort(bad_t, kind="stable")
    assert good.tolist() == [0, 1, 2, 3]
    assert bad.tolist() != good.tolist()


# ---- scan / reduce -----------------------------------------------------------

@pytest.mark.parametrize("n", SIZES)
def test_scan_reference_test_vectors(golden, n):
    x = golden[f"scan_in_{n}"]
    np.testing.assert_array_equal(oracle.scan(x, "plus", 0, False),
                                  golden[f"scan_expect_inclusive_{n}"])
    np.testing.assert_array_equal(oracle.scan(x, "plus", 0, True),
                                  golden[f"scan_expect_exclusive_{n}"])


def test_scan_projection_and_max_reference_vectors(golden):
    v = golden["proj_in"]
    np.testing.assert_array_equal(oracle.scan(v, "right", 0), v)
    np.testing.assert_array_equal(oracle.scan(v, "maximum", np.iinfo(np.int32).min),
                                  golden["max_scan_expect"])


def test_scan_closed_form_with_wrap():
    rng = np.random.default_rng(3)
    x = rng.integers(0, 2**32, size=100000, dtype=np.uint32)
    inc = np.cumsum(x, dtype=np.uint64).astype(np.uint32)  # mod 2^32
    np.testing.assert_array_equal(oracle.scan(x, "plus", 0), inc)
    exc = np.concatenate([[0], inc[:-1]]).astype(np.uint32)
    np.testing.assert_array_equal(oracle.scan(x, "plus", 0, exclusive=True), exc)
    np.testing.assert_array_equal(oracle.scan(x, "bit_xor", 0),
                                  np.bitwise_xor.accumulate(x))
    np.testing.assert_array_equal(oracle.scan(x, "minimum", 0xFFFFFFFF),
                                  np.minimum.accumulate(x))


def test_scan_exclusive_starts_at_identity():
    x = np.arange(1, 6, dtype=np.int32)
    assert oracle.scan(x, "multiplies", 1, exclusive=True).tolist() == [1, 1, 2, 6, 24]
    assert oracle.scan(x, "maximum", -100, exclusive=True).tolist() == [-100, 1, 2, 3, 4]


@pytest.mark.parametrize("n", SIZES)
def test_reduce_reference_test_vectors(golden, n):
    x = golden[f"reduce_in_{n}"]
    e = golden[f"reduce_expect_{n}"]
    assert oracle.reduce(x, "plus", 0, 7) == e[0]
    if n:
        assert oracle.reduce(x, "maximum", 0, 0) == e[1]
        assert oracle.reduce(x, "minimum", 0xFFFFFFFF, 0xFFFFFFFF) == e[2]


def test_reduce_capi_vector(golden):
    x = golden["capi_in"]
    s, m = golden["capi_expect_sum_max"]
    assert oracle.reduce(x, "plus", 0, 0) == np.int32(s)
    assert oracle.reduce(x, "maximum", np.iinfo(np.int32).min,
                         np.iinfo(np.int32).min) == m


def test_reduce_empty_returns_init_and_float_minmax():
    assert oracle.reduce(np.zeros(0, np.uint32), "plus", 0, 42) == 42
    f = (np.arange(10000) % 17 - 8).astype(np.float32)   # main.cpp:200-204
    assert oracle.reduce(f, "maximum", -np.inf, -np.inf) == 8.0
    assert oracle.reduce(f, "minimum", np.inf, np.inf) == -8.0
    with pytest.raises(ValueError):
        oracle.reduce(f, "bit_and", 0, 0)


# ---- scan_fn / reduce_fn (operators as Python functions) ---------------------

@pytest.mark.parametrize("n", SIZES)
def test_scan_fn_reference_test_vectors(golden, n):
    x = golden[f"scan_in_{n}"]
    plus = lambda a, b: a + b  # noqa: E731
    np.testing.assert_array_equal(oracle.scan_fn(x, plus, 0),
                                  golden[f"scan_expect_inclusive_{n}"])
    np.testing.assert_array_equal(oracle.scan_fn(x, plus, 0, exclusive=True),
                                  golden[f"scan_expect_exclusive_{n}"])


def test_scan_fn_projection_and_max_reference_vectors(golden):
    v = golden["proj_in"]
    np.testing.assert_array_equal(oracle.scan_fn(v, lambda a, b: b, 0), v)
    np.testing.assert_array_equal(oracle.scan_fn(v, max, np.iinfo(np.int32).min),
                                  golden["max_scan_expect"])


FNS = {"plus": lambda a, b: a + b, "multiplies": lambda a, b: a * b,
       "bit_and": lambda a, b: a & b, "bit_or": lambda a, b: a | b,
       "bit_xor": lambda a, b: a ^ b, "minimum": lambda a, b: b if b < a else a,
       "maximum": lambda a, b: b if a < b else a, "right": lambda a, b: b,
       "left": lambda a, b: a}


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32,
                                   np.uint32, np.int64, np.uint64, np.float32,
                                   np.float64])
def test_scan_fn_and_reduce_fn_match_the_c_oracle(dtype):
    """Every canonical operator, every type, wrap-around included: the
    Python restatement equals oracle.c's (pinned by the vectors above)."""
    dt = np.dtype(dtype)
    rng = np.random.default_rng(dt.itemsize * 7 + (dt.kind == "f"))
    n = 3001
    if dt.kind == "f":
        x = rng.integers(-50, 50, size=n).astype(dt)
        ops = ["plus", "minimum", "maximum", "right", "left"]
    else:
        info = np.iinfo(dt)
        x = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
        ops = list(FNS)
    for op in ops + (["multiplies"] if dt.kind == "f" else []):
        ident = 1 if op == "multiplies" else 0
        xs = x[:40] if (op == "multiplies" and dt.kind == "f") else x
        for ex in (False, True):
            np.testing.assert_array_equal(oracle.scan_fn(xs, FNS[op], ident, ex),
                                          oracle.scan(xs, op, ident, ex), err_msg=op)
        if op not in ("right", "left"):
            got = oracle.reduce_fn(xs, FNS[op], ident, 3)
            want = oracle.reduce(xs, op, ident, 3)
            assert got == want, op
    assert oracle.reduce_fn(x[:0], FNS["plus"], 0, 5) == 5


def _affine(a, b):
    """Maps x -> m*x + c mod 2^32 packed as c << 32 | m; a then b."""
    m32 = 0xFFFFFFFF
    am, ac, bm, bc = a & m32, a >> 32, b & m32, b >> 32
    return (((bm * ac + bc) & m32) << 32) | ((am * bm) & m32)


def test_scan_fn_non_commutative_order_closed_form():
    """Affine-map composition (associative, not commutative): out[i] must be
    the map 'x[0], then x[1], ..., then x[i]', checked by applying the maps
    to points one by one; a scan that swapped operands fails."""
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2**64, size=500, dtype=np.uint64, endpoint=False)
    out = oracle.scan_fn(x, _affine, 1)
    ex = oracle.scan_fn(x, _affine, 1, exclusive=True)
    assert int(ex[0]) == 1
    for t in (0, 1, 12345):
        v = t
        for i in range(x.size):
            m, c = int(x[i]) & 0xFFFFFFFF, int(x[i]) >> 32
            v = (m * v + c) & 0xFFFFFFFF
            g = int(out[i])
            assert ((g & 0xFFFFFFFF) * t + (g >> 32)) & 0xFFFFFFFF == v
    swapped = oracle.scan_fn(x, lambda a, b: _affine(b, a), 1)
    assert not np.array_equal(swapped, out)


# ---- sequence -------------------------------------------------------------------

def test_sequence_capi_closed_form():
    """capi.c:139-141: sequence(start 5, step 2) gives 5 + 2i."""
    n = 100000
    out = oracle.sequence(n, np.int32, 5, 2)
    assert out[0] == 5 and out[-1] == 5 + 2 * (n - 1)
    np.testing.assert_array_equal(out, 5 + 2 * np.arange(n, dtype=np.int64))


def test_sequence_wraps_and_rounds():
    u8 = oracle.sequence(600, np.uint8, 250, 3)
    assert u8.tolist() == [(250 + 3 * i) % 256 for i in range(600)]
    i8 = oracle.sequence(300, np.int8, -128, -1)
    assert i8.tolist() == [((-128 - i + 128) % 256) - 128 for i in range(300)]
    i64 = oracle.sequence(5, np.int64, 2**63 - 2, 1)
    assert i64.tolist() == [2**63 - 2, 2**63 - 1, -2**63, -2**63 + 1, -2**63 + 2]
    f = oracle.sequence(1 << 25, np.float32, 0.5, 1.0)
    i = np.arange(1 << 25, dtype=np.float64)
    want = np.float32(np.float32(i) * np.float32(1.0)) + np.float32(0.5)
    np.testing.assert_array_equal(f, want.astype(np.float32))
    assert f[(1 << 24) + 1] == np.float32((1 << 24) + 0.5)  # (T)i rounds to 2^24


# ---- copy_if -------------------------------------------------------------------

@pytest.mark.parametrize("n", SIZES)
def test_copy_if_reference_test_rule(golden, n):
    """main.cpp:316-329: copy_if(x, "(x % 3u) == 0u") on the histogram test's
    inputs equals the in-order list of multiples of 3."""
    x = golden[f"hist_in_{n}"]
    out, count = oracle.copy_if(x, lambda v: v % 3 == 0)
    ref = [int(v) for v in x if v % 3 == 0]
    assert count == len(ref)
    assert out.tolist() == ref


def test_copy_if_capi_vector(golden):
    """capi.c:130-131: copy_if(x > 0) keeps every positive element."""
    x = golden["capi_in"]
    out, count = oracle.copy_if(x, lambda v: v > 0)
    assert count == int((x > 0).sum())
    assert out.tolist() == [int(v) for v in x if v > 0]


def test_copy_if_edges():
    x = np.arange(10, dtype=np.int64)
    assert oracle.copy_if(x, lambda v: v < 0)[1] == 0
    out, c = oracle.copy_if(x, lambda v: v >= 0)
    assert c == 10 and out.tolist() == list(range(10))
    assert oracle.copy_if(x[:0], lambda v: v > 0)[1] == 0
    f = np.array([-0.0, 0.0, np.nan, 1.5, -2.0], dtype=np.float32)
    out, c = oracle.copy_if(f, lambda v: v <= 0)
    assert c == 3 and np.signbit(out[0]) and not np.signbit(out[1]) and out[2] == -2.0


# ---- histogram ---------------------------------------------------------------

@pytest.mark.parametrize("n", SIZES)
def test_histogram_reference_test_vectors(golden, n):
    x = golden[f"hist_in_{n}"]
    np.testing.assert_array_equal(oracle.histogram_even(x, 10, 0.0, 1000.0),
                                  golden[f"hist_expect_{n}"])


@pytest.mark.parametrize("dtype", [np.uint32, np.float32, np.int64])
@pytest.mark.parametrize("descending", [False, True])
def test_segmented_sort_matches_per_segment_numpy(dtype, descending):
    rng = np.random.default_rng(21)
    off = oracle.lognormal_segments(300, np.log(40), 1.5, 0, 5000, seed=4)
    n = int(off[-1])
    if np.dtype(dtype).kind == "f":
        keys = rng.standard_normal(n).astype(dtype)
        keys[::7] = 0.0
        keys[::11] = -0.0
    else:
        keys = rng.integers(-50, 50, size=n).astype(dtype)
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.segmented_sort(keys, vals, off, descending, threads=4)
    for i in range(off.size - 1):
        b, e = int(off[i]), int(off[i + 1])
        perm = np.argsort(oracle.transformed_keys(keys[b:e], descending), kind="stable")
        np.testing.assert_array_equal(v[b:e], vals[b:e][perm])
        np.testing.assert_array_equal(k[b:e].view(np.uint8), keys[b:e][perm].view(np.uint8))


def test_lognormal_segments_shape():
    off = oracle.lognormal_segments(65536, np.log(512), 1.5, 1, 1 << 20, seed=1)
    sizes = np.diff(off)
    assert sizes.min() >= 1 and sizes.max() <= (1 << 20)
    assert 400 <= np.median(sizes) <= 650          # median exp(mu) = 512
    assert 2**25 < off[-1] < 2**27                  # ~ 65536 * 512 * e^1.125


def test_histogram_brute_force_float_bins():
    rng = np.random.default_rng(9)
    x = (rng.standard_normal(50000) * 3).astype(np.float32)
    lo, up, bins = np.float32(-4.0), np.float32(5.5), 37
    scale = np.float32(bins) / (up - lo)
    expect = np.zeros(bins, np.uint32)
    for v in x[(x >= lo) & (x < up)]:
        b = int(np.float32(v - lo) * scale)
        expect[min(b, bins - 1)] += 1
    np.testing.assert_array_equal(oracle.histogram_even(x, bins, -4.0, 5.5), expect)
    with pytest.raises(ValueError):
        oracle.histogram_even(x, 4, 1.0, 1.0)
