"""The tuned-table file format on the CPU: tests/cpp/tuning_file_test.cpp
built with g++ against opencl-sdk_amd/csrc/tuning_file.cpp (no HIP), also
under ThreadSanitizer for the concurrent writers and readers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "cpp", "tuning_file_test.cpp"),
       os.path.join(ROOT, "opencl-sdk_amd", "csrc", "tuning_file.cpp")]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("san", [None, "thread", "address"])
def test_tuning_file(tmp_path, san):
    exe = str(tmp_path / "tuning_file_test")
    extra = [f"-fsanitize={san}"] if san else []
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", *extra, "-o", exe, *SRC],
                   check=True)
    work = tmp_path / "cache"
    work.mkdir()
    r = subprocess.run([exe, str(work)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0 and "FATAL: ThreadSanitizer" in r.stderr:
        pytest.skip("sanitizer runtime unavailable in this environment")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
    assert "WARNING: ThreadSanitizer" not in r.stderr
