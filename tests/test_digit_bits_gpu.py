"""Digit widths of the segmented sort's LDS passes, on the GPU.

A segment sorted in LDS whose keys differ in their low k bits takes
ceil(k / w) passes, w the width its size bucket's table entry names (1..8
bits), the k bits split evenly over the passes. The autotuner stores the
widths it picks beside the classes; tables with every width from 1 to 8
bits, planted in a context's tuning cache, sort every size class and key
type bit-exactly against the oracle's per-segment restatement of the
reference sort (Algorithm.cpp:453-527 sorts one segment per call)."""
import glob
import os

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import TYPE_OF, dev, host, say
from test_tuning_cache_gpu import NBUCKET, LARGE, write_table
from test_segmented_gpu import boundary_offsets

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tuned(tmp_path_factory):
    """Tunes u32 pairs under a fresh cache directory: (cache, header, arch
    suffix of the file names, classes, widths)."""
    cache = str(tmp_path_factory.mktemp("digit_bits"))
    say("digit bits: tune segmented u32 pairs under", cache)
    a = clc_hip.Context(cache_dir=cache)
    try:
        a.segmented_sort_tune(clc_hip.U32, True)
        classes = a.segmented_sort_table(clc_hip.U32, True)
        widths = a.segmented_sort_digit_bits(clc_hip.U32, True)
    finally:
        a.close()
    files = glob.glob(os.path.join(cache, f"clc_segsort_t{clc_hip.U32}_kv_*.txt"))
    assert len(files) == 1
    header, line, _ = open(files[0]).read().splitlines()
    arch = os.path.basename(files[0])[len(f"clc_segsort_t{clc_hip.U32}_kv_"):-len(".txt")]
    return cache, header, arch, classes, widths, [int(x) for x in line.split()]


def test_tuner_stores_widths(tuned):
    _, _, _, classes, widths, stored = tuned
    assert stored == classes + widths
    for c, w in zip(classes, widths):
        assert 4 <= w <= 8
        if c == LARGE:
            assert w == 8  # the onesweep path has no LDS passes to tune


def planted_widths(shift):
    return [1 + (b + shift) % 8 for b in range(NBUCKET)]


@pytest.mark.parametrize("dtype,with_values", [
    (np.uint32, True), (np.uint32, False), (np.float32, True),
    (np.uint64, True), (np.uint64, False), (np.float64, True)])
def test_planted_widths_sort_exactly(tuned, dtype, with_values):
    cache, header, arch, _, _, _ = tuned
    dt = np.dtype(dtype)
    code = TYPE_OF[dt]
    probe = clc_hip.Context(cache_dir=None)
    try:
        classes = probe.segmented_sort_table(code, with_values)
    finally:
        probe.close()
    for shift in range(0, 8, 3):
        widths = planted_widths(shift)
        path = os.path.join(cache, f"clc_segsort_t{code}_{'kv' if with_values else 'k'}_{arch}.txt")
        write_table(path, header, classes + widths)
        ctx = clc_hip.Context(cache_dir=cache)
        try:
            assert ctx.segmented_sort_digit_bits(code, with_values) == widths
            assert ctx.segmented_sort_table(code, with_values) == classes
            off = boundary_offsets()
            n = int(off[-1])
            rng = np.random.default_rng(100 + shift)
            if dt.kind == "f":
                keys = (rng.standard_normal(n) * 1000).astype(dt)
                keys[::11] = -0.0
                keys[::13] = 0.0
            elif dt.itemsize == 8:
                # Ranges of 20, 48 and 64 bits by thirds: the packed-word
                # path's one and two runs of passes.
                keys = oracle.pcg32_u64(shift, 64, n, bits=64)
                keys[: n // 3] &= np.uint64((1 << 20) - 1)
                keys[n // 3: 2 * n // 3] &= np.uint64((1 << 48) - 1)
            else:
                keys = oracle.pcg32_u32(shift, 7, n)
                keys[: n // 2] &= np.uint32((1 << 17) - 1)  # 17 bits: uneven splits
            keys[::7] = keys[3]
            vals = np.arange(n, dtype=np.uint32) if with_values else None
            say("digit bits:", dt.name, "values" if with_values else "keys only",
                "widths from", widths[:9])
            kd = dev(keys)
            vd = None if vals is None else dev(vals)
            ctx.segmented_sort(kd, vd, dev(off), key_type=code)
            ctx.finish()
            ek, ev = oracle.segmented_sort(keys, vals, off, threads=8)
            np.testing.assert_array_equal(host(kd, dt).view(np.uint8), ek.view(np.uint8))
            if with_values:
                np.testing.assert_array_equal(host(vd, np.uint32), ev)
        finally:
            ctx.close()


def test_out_of_range_widths_ignored(tuned):
    cache, header, arch, classes, widths, _ = tuned
    path = os.path.join(cache, f"clc_segsort_t{clc_hip.U32}_kv_{arch}.txt")
    default = clc_hip.Context(cache_dir=None)
    try:
        base_c = default.segmented_sort_table(clc_hip.U32, True)
        base_w = default.segmented_sort_digit_bits(clc_hip.U32, True)
    finally:
        default.close()
    assert base_w == [8] * NBUCKET
    for bad in (0, 9):
        write_table(path, header, classes + [bad] * NBUCKET)
        ctx = clc_hip.Context(cache_dir=cache)
        try:
            assert ctx.segmented_sort_table(clc_hip.U32, True) == base_c
            assert ctx.segmented_sort_digit_bits(clc_hip.U32, True) == base_w
        finally:
            ctx.close()
    # A table of classes alone (the earlier form) keeps 8-bit digits.
    write_table(path, header, classes)
    ctx = clc_hip.Context(cache_dir=cache)
    try:
        assert ctx.segmented_sort_table(clc_hip.U32, True) == classes
        assert ctx.segmented_sort_digit_bits(clc_hip.U32, True) == [8] * NBUCKET
    finally:
        ctx.close()
