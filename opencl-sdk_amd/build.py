"""Builds libclc_hip.so (the C-ABI library) for gfx950, in-tree.

    python opencl-sdk_amd/build.py [--jobs N] [--incremental]

Every source under csrc/ is compiled by hipcc for gfx950 into _build/ and
linked into libclc_hip.so beside this file. hipcc cross-compiles without a
GPU. --incremental skips sources whose object is newer than the source and
every file it includes (hipcc writes the list beside each object).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libclc_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-result",
    # Keep the compiler from forming packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Headers compiled again at run time by hipRTC (operator expressions of
# clc_scan / clc_reduce); their text is embedded in the library.
RTC_HEADERS = ["device_common.hpp", "scan_kernels.hpp", "elementwise_kernels.hpp"]
RTC_DELIM = "CLCHDR"  # raw-string delimiters hold at most 16 characters


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip"))
                  + glob.glob(os.path.join(CSRC, "*.cpp")))


def embedded_headers_source() -> str:
    """Writes _build/rtc_embedded.cpp, defining the header names and texts
    rtc.cpp hands to hipRTC, and returns its path (rewritten only when the
    text changes, so incremental builds stay incremental)."""
    names, texts = [], []
    for h in RTC_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            text = f.read()
        if f"){RTC_DELIM}\"" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        names.append(f'"{h}"')
        texts.append(f'R"{RTC_DELIM}({text}){RTC_DELIM}"')
    body = ("// Generated by build.py: headers for run-time compilation.\n"
            "namespace clc {\nnamespace rtc {\n"
            f"extern const int embedded_header_count = {len(names)};\n"
            "extern const char* const embedded_header_names[] = {"
            + ", ".join(names) + "};\n"
            "extern const char* const embedded_header_texts[] = {\n"
            + ",\n".join(texts) + "};\n} // namespace rtc\n} // namespace clc\n")
    path = os.path.join(BUILD, "rtc_embedded.cpp")
    old = None
    if os.path.exists(path):
        with open(path) as f:
            old = f.read()
    if old != body:
        with open(path, "w") as f:
            f.write(body)
    return path


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", src, "-o", obj, "-MMD", "-MF", obj + ".d"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def _deps(obj: str) -> list[str] | None:
    """Files the object was compiled from (its make-style .d file), or None
    when the list is missing."""
    try:
        with open(obj + ".d") as f:
            text = f.read().replace("\\\n", " ")
    except OSError:
        return None
    _, _, rest = text.partition(":")
    return rest.split()


def _fresh(src: str) -> bool:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    deps = _deps(obj)
    if not os.path.exists(obj) or deps is None:
        return False
    t = os.path.getmtime(obj)
    try:
        return all(os.path.getmtime(d) < t for d in [src, *deps])
    except OSError:
        return False


def build(jobs: int = 8, verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every source (or, with incremental, those older than their
    object or than any header under csrc/) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    extra: list[str] = []
    srcs = sources() + [embedded_headers_source()]

    def one(s):
        if incremental and _fresh(s):
            return os.path.join(BUILD, os.path.basename(s) + ".o")
        return _compile(s, extra)

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources older than their object or any file "
                         "they include")
    args = ap.parse_args()
    build(jobs=args.jobs, verbose=True, incremental=args.incremental)


if __name__ == "__main__":
    sys.exit(main())
