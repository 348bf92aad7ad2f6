"""The benchmark's input generators on the CPU: bench.pcg32_stream (the
reference's pcg32, Random.c:13-50, computed by state doubling in numpy)
must reproduce the oracle's sequential pcg32 exactly, and the published
pcg32 demo sequence for srandom(42, 54)."""
import numpy as np
import pytest

import bench
import oracle


def test_published_pcg32_demo_sequence():
    assert [int(x) for x in bench.pcg32_stream(42, 54, 6)] == [
        0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


@pytest.mark.parametrize("seed", [(42, 54), (42, 55), (42, 56), (42, 57), (7, 11),
                                  (0, 0), (2**64 - 1, 2**63 + 5)])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 1000, (1 << 16) + 7])
def test_stream_equals_sequential_pcg32(seed, n):
    np.testing.assert_array_equal(bench.pcg32_stream(*seed, n), oracle.pcg32_u32(*seed, n))


def test_u64_keys_construction():
    raw = bench.pcg32_stream(42, 56, 2000).astype(np.uint64)
    k = ((raw[0::2] << np.uint64(32)) | raw[1::2]) & np.uint64((1 << 48) - 1)
    np.testing.assert_array_equal(k, oracle.pcg32_u64(42, 56, 1000, bits=48))
