"""The autotuners' on-disk cache on the GPU (tuned tables in a context's
cache directory; the reference's program cache keeps its binaries the same
way, lib/src/Compute/ProgramCache.cpp:163-219).

Tune under a cache directory, then open new contexts on it: a fresh
context reads the stored table instead of the default; a different valid
table planted in the file is what the next context reports (so the table
really comes from disk, not from re-tuning); a damaged or foreign file is
ignored (the default table comes back); sorts through each context stay
bit-exact against the oracle."""
import glob
import os

import numpy as np
import pytest

import clc_hip
import oracle
from gpu_helpers import dev, host, say

pytestmark = pytest.mark.gpu

NBUCKET = 32
LARGE = 7


def fnv1a(text: str) -> int:
    h = 0xcbf29ce484222325
    for c in text.encode():
        h = ((h ^ c) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def write_table(path, header, values):
    line = " ".join(str(v) for v in values)
    with open(path, "w") as f:
        f.write(f"{header}\n{line}\nend {len(values)} {fnv1a(line):016x}\n")


def segmented_check(ctx):
    off = np.array([0, 3000, 3000, 9000, 70000, 70001, 200000], dtype=np.uint64)
    n = int(off[-1])
    keys = oracle.pcg32_u32(17, 3, n)
    vals = np.arange(n, dtype=np.uint32)
    kd, vd = dev(keys), dev(vals)
    ctx.segmented_sort(kd, vd, dev(off), key_type=clc_hip.U32)
    ctx.finish()
    ek, ev = oracle.segmented_sort(keys, vals, off, threads=8)
    np.testing.assert_array_equal(host(kd, np.uint32), ek)
    np.testing.assert_array_equal(host(vd, np.uint32), ev)


def test_segmented_table_round_trip_through_disk(tmp_path):
    import torch
    cache = str(tmp_path)
    default = clc_hip.Context(cache_dir=None)
    try:
        base = default.segmented_sort_table(clc_hip.U32, True)
    finally:
        default.close()

    say("tuning cache: tune segmented u32 pairs under", cache)
    a = clc_hip.Context(cache_dir=cache)
    try:
        a.segmented_sort_tune(clc_hip.U32, True)
        tuned = a.segmented_sort_table(clc_hip.U32, True)
        tuned_bits = a.segmented_sort_digit_bits(clc_hip.U32, True)
    finally:
        a.close()
    files = glob.glob(os.path.join(cache, "clc_segsort_t5_kv_*.txt"))
    assert len(files) == 1
    assert not glob.glob(os.path.join(cache, "*.tmp.*"))
    path = files[0]
    header, line, end = open(path).read().splitlines()
    assert [int(x) for x in line.split()] == tuned + tuned_bits  # classes, then digit widths
    assert header.startswith("clc-tuning-v2 gfx950")

    say("tuning cache: a new context reads the stored table")
    b = clc_hip.Context(cache_dir=cache)
    try:
        assert b.segmented_sort_table(clc_hip.U32, True) == tuned
        assert b.segmented_sort_digit_bits(clc_hip.U32, True) == tuned_bits
    finally:
        b.close()

    say("tuning cache: a planted table is what the next context uses")
    # Segments of 3..32 keys in the 1024 x 16 configuration: valid (it holds
    # them), but never what the tuner would pick for them.
    planted = [6 if 2 <= i <= 5 else c for i, c in enumerate(base)]
    assert planted != tuned and planted != base
    write_table(path, header, planted)
    c = clc_hip.Context(cache_dir=cache)
    try:
        assert c.segmented_sort_table(clc_hip.U32, True) == planted
        segmented_check(c)
    finally:
        c.close()

    for label, text in (("damaged", open(path).read().replace("end ", "end 1")),
                        ("foreign", open(path).read().replace("gfx950", "gfx900", 1)),
                        ("infeasible", None)):
        say("tuning cache:", label, "file ignored")
        if text is None:
            write_table(path, header, [99] * NBUCKET)  # no such class
        else:
            with open(path, "w") as f:
                f.write(text)
        d = clc_hip.Context(cache_dir=cache)
        try:
            assert d.segmented_sort_table(clc_hip.U32, True) == base
            segmented_check(d)
        finally:
            d.close()
    torch.cuda.synchronize()


def test_sort_shape_stored_and_reused(tmp_path):
    cache = str(tmp_path)
    say("tuning cache: onesweep shape for u32 pairs")
    a = clc_hip.Context(cache_dir=cache)
    try:
        shape = a.sort_tune(clc_hip.U32, True)
    finally:
        a.close()
    files = glob.glob(os.path.join(cache, "clc_sort_k32_kv_*.txt"))
    assert len(files) == 1
    header, line, _ = open(files[0]).read().splitlines()
    assert [int(x) for x in line.split()] == [shape]
    # Every stored shape gives the oracle's output through a new context.
    for planted in range(11):
        write_table(files[0], header, [planted])
        b = clc_hip.Context(cache_dir=cache)
        try:
            n = 100003
            keys = oracle.pcg32_u32(planted, 9, n)
            vals = np.arange(n, dtype=np.uint32)
            kd, vd = dev(keys), dev(vals)
            b.sort(kd, vd, key_type=clc_hip.U32)
            b.finish()
            ek, ev = oracle.radix_sort(keys, vals, threads=8)
            np.testing.assert_array_equal(host(kd, np.uint32), ek)
            np.testing.assert_array_equal(host(vd, np.uint32), ev)
        finally:
            b.close()
