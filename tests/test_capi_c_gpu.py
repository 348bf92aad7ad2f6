"""tests/cpp/capi_test.c on the GPU: the C interface driven from C with
library-owned buffers (clc_buffer_*), following the reference's C test
(lib/test/Compute/capi.c). Compiled with gcc as C99 against
include/compute_c.h and the built library (a second's work), then run."""
import os
import shutil
import subprocess

import pytest

from gpu_helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "opencl-sdk_amd")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_capi_c(tmp_path):
    exe = str(tmp_path / "capi_test")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-O1",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "capi_test.c"), "-o", exe,
                    "-L", LIBDIR, "-lclc_hip", "-lm", f"-Wl,-rpath,{LIBDIR}"], check=True)
    say("capi_test.c")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 failures" in r.stdout
