"""Python face of the CPU oracle (oracle.c) plus the seeded input generators.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The shipped library never
imports it. Every function names the reference code it restates; oracle.c
holds the algorithms, this file only marshals numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# clc_type codes (reference lib/include/CL/Compute/compute_c.h:44-56).
TYPE_CODES = {
    np.dtype(np.int8): 0, np.dtype(np.uint8): 1,
    np.dtype(np.int16): 2, np.dtype(np.uint16): 3,
    np.dtype(np.int32): 4, np.dtype(np.uint32): 5,
    np.dtype(np.int64): 6, np.dtype(np.uint64): 7,
    np.dtype(np.float32): 8, np.dtype(np.float64): 9,
}

# Operator codes of oracle.c: the canonical ops:: strings of the reference
# (lib/include/CL/Compute/Algorithm.hpp:53-65) and the projections "b"/"a".
OPS = {"plus": 0, "multiplies": 1, "bit_and": 2, "bit_or": 3, "bit_xor": 4,
       "minimum": 5, "maximum": 6, "right": 7, "left": 8}

_lib = None


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, sz, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64
        L.oracle_pcg32_u32.argtypes = [u64, u64, u64, vp, sz]
        L.oracle_pcg32_u32.restype = None
        L.oracle_ints_range.argtypes = [u64, u64, vp, sz, ctypes.c_int32,
                                        ctypes.c_int32]
        L.oracle_ints_range.restype = None
        L.oracle_radix_sort.argtypes = [ctypes.c_int, vp, vp, sz, ctypes.c_int,
                                        ctypes.c_int]
        L.oracle_radix_sort.restype = ctypes.c_int
        L.oracle_scan.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, sz,
                                  ctypes.c_int]
        L.oracle_scan.restype = ctypes.c_int
        L.oracle_reduce.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, sz,
                                    vp]
        L.oracle_reduce.restype = ctypes.c_int
        L.oracle_histogram_even.argtypes = [ctypes.c_int, vp, sz, vp, sz,
                                            ctypes.c_double, ctypes.c_double]
        L.oracle_histogram_even.restype = ctypes.c_int
        L.oracle_segmented_sort.argtypes = [ctypes.c_int, vp, vp, vp, sz, ctypes.c_int,
                                            ctypes.c_int]
        L.oracle_segmented_sort.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---- generators (reference lib/src/SDK/Random.c) -------------------------

def pcg32_u32(initstate: int, initseq: int, n: int, skip: int = 0) -> np.ndarray:
    """n raw pcg32_random_r outputs after pcg32_srandom_r(initstate, initseq)
    (Random.c:13-22, :43-50), starting `skip` draws into the stream."""
    out = np.empty(n, dtype=np.uint32)
    lib().oracle_pcg32_u32(initstate, initseq, skip, _ptr(out), n)
    return out


def ints_range(initstate: int, initseq: int, n: int, low: int, hi: int) -> np.ndarray:
    """cl_sdk_fill_with_random_ints_range (Random.c:73-102)."""
    out = np.empty(n, dtype=np.int32)
    lib().oracle_ints_range(initstate, initseq, _ptr(out), n, low, hi)
    return out


def pcg32_u64(initstate: int, initseq: int, n: int, bits: int = 64) -> np.ndarray:
    """64-bit keys from two pcg32 draws, (hi << 32 | lo), masked to `bits`."""
    raw = pcg32_u32(initstate, initseq, 2 * n).astype(np.uint64)
    k = (raw[0::2] << np.uint64(32)) | raw[1::2]
    if bits < 64:
        k &= np.uint64((1 << bits) - 1)
    return k


# ---- key transform (reference lib/src/Compute/Algorithm.cpp:459-469) -------

def transformed_keys(keys: np.ndarray, descending: bool = False) -> np.ndarray:
    """The unsigned key whose ascending order the reference sort follows:
    uint k; int k ^ signbit; float (k & sign) ? ~k : k | sign; descending ~.
    Extended to 64-bit keys with the same formulas."""
    dt = keys.dtype
    if dt.itemsize == 4:
        u = keys.view(np.uint32).copy()
        top = np.uint32(0x80000000)
    elif dt.itemsize == 8:
        u = keys.view(np.uint64).copy()
        top = np.uint64(0x8000000000000000)
    else:
        raise ValueError("keys must be 32- or 64-bit")
    if dt.kind == "i":
        u ^= top
    elif dt.kind == "f":
        neg = (u & top) != 0
        u = np.where(neg, ~u, u | top)
    if descending:
        u = ~u
    return u


# ---- algorithms -----------------------------------------------------------

def radix_sort(keys: np.ndarray, values: np.ndarray | None = None,
               descending: bool = False, threads: int = 1):
    """Stable LSD radix sort restated from Algorithm.cpp:129-213, :453-527.
    Returns sorted copies (keys, values-or-None)."""
    k = np.ascontiguousarray(keys).copy()
    v = None if values is None else np.ascontiguousarray(values, dtype=np.uint32).copy()
    code = TYPE_CODES[k.dtype]
    rc = lib().oracle_radix_sort(code, _ptr(k), _ptr(v), k.size,
                                 1 if descending else 0, threads)
    if rc == -1:
        raise ValueError("oracle_radix_sort: key type must be 32- or 64-bit")
    if rc != 0:
        raise MemoryError("oracle_radix_sort: host allocation failed")
    return k, v


def segmented_sort(keys: np.ndarray, values, offsets: np.ndarray,
                   descending: bool = False, threads: int = 1):
    """Each segment [offsets[i], offsets[i+1]) sorted by the restated LSD
    sort on its own. Returns sorted copies (keys, values-or-None)."""
    k = np.ascontiguousarray(keys).copy()
    v = None if values is None else np.ascontiguousarray(values, dtype=np.uint32).copy()
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    rc = lib().oracle_segmented_sort(TYPE_CODES[k.dtype], _ptr(k), _ptr(v), _ptr(off),
                                     off.size - 1, 1 if descending else 0, threads)
    if rc != 0:
        raise ValueError("oracle_segmented_sort failed")
    return k, v


def lognormal_segments(nseg: int, mu: float, sigma: float, lo: int, hi: int,
                       seed: int) -> np.ndarray:
    """Offsets (nseg + 1, uint64) of segments with sizes
    round(exp(N(mu, sigma))) clamped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.rint(np.exp(rng.normal(mu, sigma, nseg))), lo, hi).astype(np.uint64)
    off = np.zeros(nseg + 1, dtype=np.uint64)
    np.cumsum(sizes, out=off[1:])
    return off


def zipf_keys(n: int, s: float, hi: int, seed: int) -> np.ndarray:
    """Zipf(s) ranks truncated to [1, hi] by rejection (heavy duplicates)."""
    rng = np.random.default_rng(seed)
    out = rng.zipf(s, n).astype(np.uint64)
    bad = out > hi
    while bad.any():
        out[bad] = rng.zipf(s, int(bad.sum())).astype(np.uint64)
        bad = out > hi
    return out


def lognormal_keys(n: int, mu: float, sigma: float, seed: int) -> np.ndarray:
    """round(lognormal(mu, sigma)) as uint64."""
    rng = np.random.default_rng(seed)
    return np.rint(rng.lognormal(mu, sigma, n)).astype(np.uint64)


def gaussian_mixture_keys(n: int, k: int, span: float, sd: float, seed: int) -> np.ndarray:
    """k Gaussians with means uniform in [0, span) and standard deviation sd,
    clamped to [0, 2^64 - 1], as uint64."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.0, span, k)
    x = means[rng.integers(0, k, n)] + rng.normal(0.0, sd, n)
    return np.clip(x, 0.0, 2.0**64 - 2.0**12).astype(np.uint64)


def scan(x: np.ndarray, op: str = "plus", identity=0, exclusive: bool = False) -> np.ndarray:
    """Sequential left-to-right scan (Algorithm.cpp:76-127, :280-318)."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    ident = np.array([identity], dtype=x.dtype)
    rc = lib().oracle_scan(TYPE_CODES[x.dtype], OPS[op], _ptr(ident), _ptr(x),
                           _ptr(out), x.size, 1 if exclusive else 0)
    if rc != 0:
        raise ValueError(f"oracle_scan: operator {op} undefined for {x.dtype}")
    return out


def reduce(x: np.ndarray, op: str = "plus", identity=0, init=0):
    """op(init, fold(identity, x)) (Algorithm.cpp:58-74, :400-442)."""
    x = np.ascontiguousarray(x)
    ident = np.array([identity], dtype=x.dtype)
    ini = np.array([init], dtype=x.dtype)
    res = np.zeros(1, dtype=x.dtype)
    rc = lib().oracle_reduce(TYPE_CODES[x.dtype], OPS[op], _ptr(ident),
                             _ptr(ini), _ptr(x), x.size, _ptr(res))
    if rc != 0:
        raise ValueError(f"oracle_reduce: operator {op} undefined for {x.dtype}")
    return res[0]


def _element_converter(dtype):
    """Python number -> the value C's conversion to dtype gives: integers
    wrap modulo 2^bits (two's complement), floats round to the type."""
    dt = np.dtype(dtype)
    if dt.kind in "iu":
        bits = dt.itemsize * 8
        mask = (1 << bits) - 1
        signed = dt.kind == "i"

        def conv(v):
            v = int(v) & mask
            return v - (1 << bits) if signed and v >> (bits - 1) else v
        return conv
    if dt == np.float32:
        return lambda v: float(np.float32(v))
    return float


def scan_fn(x: np.ndarray, fn, identity, exclusive: bool = False) -> np.ndarray:
    """The reference's scan (running = IDENTITY; next = OP(running, x[i]);
    out[i] = running if exclusive else next -- Algorithm.cpp:85-104) with
    the operator given as a Python function of two Python numbers, for
    operator expressions the library compiles at run time. Each result is
    converted to x's dtype as C converts the expression's value (T)(OP).
    A plain loop: small n only."""
    conv = _element_converter(x.dtype)
    run = conv(identity)
    res = []
    for v in x.tolist():
        nxt = conv(fn(run, v))
        res.append(run if exclusive else nxt)
        run = nxt
    return np.array(res, dtype=x.dtype)


def reduce_fn(x: np.ndarray, fn, identity, init):
    """OP(init, fold) with the fold seeded by IDENTITY (Algorithm.cpp:58-74,
    :400-442), the operator a Python function as in scan_fn. n == 0 gives
    init."""
    conv = _element_converter(x.dtype)
    if x.size == 0:
        return conv(init)
    acc = conv(identity)
    for v in x.tolist():
        acc = conv(fn(acc, v))
    return conv(fn(conv(init), acc))


def sequence(n: int, dtype, start, step) -> np.ndarray:
    """The reference's sequence: out[i] = start + (T)i * step
    (Algorithm.cpp:23-27, :345-359). Integers wrap modulo 2^bits of T (the
    OpenCL C conversion of the promoted result); floating point rounds the
    product, then the sum."""
    dt = np.dtype(dtype)
    i = np.arange(n, dtype=np.uint64)
    if dt.kind == "f":
        ti = i.astype(np.uint32).astype(dt)  # (T)i from the uint index
        return (ti * dt.type(step)).astype(dt) + dt.type(start)
    bits = dt.itemsize * 8
    mask = np.uint64((1 << bits) - 1)
    s = np.uint64(int(start) & int(mask))
    d = np.uint64(int(step) & int(mask))
    ti = i & mask
    with np.errstate(over="ignore"):
        v = (s + ti * d) & mask
    return v.astype(dt.newbyteorder("=")) if dt.kind == "u" else v.astype(
        np.dtype(f"u{dt.itemsize}")).view(dt)


def copy_if(x: np.ndarray, predicate) -> tuple[np.ndarray, int]:
    """The reference's copy_if (Algorithm.cpp:529-567): flags[i] =
    predicate(x[i]) ? 1 : 0, positions = exclusive scan of the flags (with
    uint32 '+'), out[positions[i]] = x[i] for every flagged i, count =
    positions[n-1] + flags[n-1]. predicate maps a numpy array to a boolean
    array. Returns (out[:count], count)."""
    x = np.ascontiguousarray(x)
    n = x.size
    if n == 0:
        return x[:0].copy(), 0
    flags = np.asarray(predicate(x), dtype=bool).astype(np.uint32)
    positions = scan(flags, "plus", 0, exclusive=True)
    count = int(positions[-1]) + int(flags[-1])
    out = np.zeros(n, dtype=x.dtype)
    sel = flags.astype(bool)
    out[positions[sel]] = x[sel]
    return out[:count], count


def histogram_even(x: np.ndarray, bins: int, lower: float, upper: float) -> np.ndarray:
    """Even-width histogram with the reference's float binning
    (Algorithm.cpp:240-273, :569-615)."""
    x = np.ascontiguousarray(x)
    h = np.zeros(max(bins, 1), dtype=np.uint32)
    rc = lib().oracle_histogram_even(TYPE_CODES[x.dtype], _ptr(x), x.size,
                                     _ptr(h), bins, lower, upper)
    if rc != 0:
        raise ValueError("oracle_histogram_even: lower must be below upper")
    return h[:bins]
