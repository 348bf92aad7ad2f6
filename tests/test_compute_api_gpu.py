"""tests/cpp/compute_api_test.cpp on the GPU: the C++ algorithm API
(include/compute.hpp) driven the way the reference's C++ tests drive theirs
(lib/test/Compute/main.cpp:56-89, 182-337). Compiled with g++ against the
header and the built library (a few seconds), then run."""
import os
import shutil
import subprocess

import pytest

from gpu_helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "opencl-sdk_amd")


def compile_cmd(src, exe):
    return ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-O1",
            "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
            "-I", "/opt/rocm/include", src, "-o", exe, "-L", LIBDIR, "-lclc_hip",
            "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{LIBDIR}",
            "-Wl,-rpath,/opt/rocm/lib"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_compute_api_cxx(tmp_path):
    exe = str(tmp_path / "compute_api_test")
    subprocess.run(compile_cmd(os.path.join(ROOT, "tests", "cpp", "compute_api_test.cpp"),
                               exe), check=True)
    say("compute_api_test.cpp")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 failures" in r.stdout
