"""Regenerates reference_tests.npz: the inputs and expected results of the
reference's own Compute tests (lib/test/Compute/main.cpp, capi.c).

Compiles gen_golden.cpp with g++ (libstdc++, whose std::mt19937 and
std::uniform_int_distribution define the reference tests' input streams),
runs it into a temporary directory and packs the raw files. The expected
outputs come from std::stable_sort / std::sort / running sums, exactly the
host references of those tests; the reference library is not run.

    python tests/golden/make_golden.py
"""
import glob
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

DTYPES = {
    "stable_keys_u8": np.uint8,
    "ints_in": np.int32, "ints_expect": np.int32,
    "floats_in": np.float32, "floats_expect_desc": np.float32,
    "proj_in": np.int32, "max_scan_expect": np.int32,
    "capi_in": np.int32, "capi_expect_sum_max": np.int64,
}


def main() -> None:
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "gen_golden")
        subprocess.run(["g++", "-O2", "-std=c++14", "-o", exe,
                        os.path.join(HERE, "gen_golden.cpp")], check=True)
        out = os.path.join(tmp, "out")
        os.mkdir(out)
        subprocess.run([exe, out], check=True)
        arrays = {}
        for path in sorted(glob.glob(os.path.join(out, "*.bin"))):
            name = os.path.basename(path)[:-4]
            arrays[name] = np.fromfile(path, dtype=DTYPES.get(name, np.uint32))
    np.savez_compressed(os.path.join(HERE, "reference_tests.npz"), **arrays)
    print(f"wrote {len(arrays)} arrays")


if __name__ == "__main__":
    main()
