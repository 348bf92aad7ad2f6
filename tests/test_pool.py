"""The per-GPU scratch pool on the CPU: tests/cpp/stream_pool_test.cpp built
with g++ against opencl-sdk_amd/csrc/stream_pool.cpp and a fake device link
(no HIP involved), plus ThreadSanitizer and AddressSanitizer builds of the
same test for the pool's locking and its block bookkeeping."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "cpp", "stream_pool_test.cpp"),
       os.path.join(ROOT, "opencl-sdk_amd", "csrc", "stream_pool.cpp")]


def build_and_run(tmp_path, extra):
    exe = str(tmp_path / "stream_pool_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", *extra, "-o", exe, *SRC],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    return r


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pool_resource(tmp_path):
    r = build_and_run(tmp_path, [])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("san", ["thread", "address"])
def test_pool_resource_sanitized(tmp_path, san):
    r = build_and_run(tmp_path, [f"-fsanitize={san}"])
    if r.returncode != 0 and "FATAL: ThreadSanitizer" in r.stderr:
        pytest.skip("sanitizer runtime unavailable in this environment")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr
