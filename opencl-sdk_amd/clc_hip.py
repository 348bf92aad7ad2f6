"""ctypes binding of libclc_hip.so (include/compute_c.h) for Python callers,
the tests and the benchmark.

Device memory and streams come from torch tensors (data_ptr() and
torch.cuda streams); every computation runs in the HIP library through the
C ABI. Loading fails loudly when the library has not been built: there is
no fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclc_hip.so")

# clc_status (compute_c.h)
SUCCESS = 0
ERROR_OPENCL = 1
ERROR_INVALID_ARGUMENT = 2
ERROR_BUILD = 3
ERROR_NO_DEVICE = 4
ERROR_OUT_OF_MEMORY = 5
ERROR_UNSUPPORTED = 6
ERROR_INTERNAL = 7

# clc_type (compute_c.h)
I8, U8, I16, U16, I32, U32, I64, U64, F32, F64 = range(10)

DEFAULT_DEVICE = ctypes.c_size_t(-1).value


class ClcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"clc status {status}: {message}")
        self.status = status
        self.message = message


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("platform", ctypes.c_char * 256),
                ("is_gpu", ctypes.c_int), ("has_fp64", ctypes.c_int),
                ("global_memory", ctypes.c_ulonglong),
                ("max_work_group_size", ctypes.c_size_t)]


class MemoryStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_size_t) for n in
                ("current_bytes", "peak_bytes", "total_bytes",
                 "current_allocations", "total_allocations", "cached_bytes",
                 "reused", "cross_stream", "stream_waits")]


class ProgramStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_size_t) for n in ("memory_hits", "disk_hits", "builds")]


class KernelTiming(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("launches", ctypes.c_size_t),
                ("total_ms", ctypes.c_double), ("min_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double)]


class LearnedStats(ctypes.Structure):
    _fields_ = [("used_model", ctypes.c_int), ("num_buckets", ctypes.c_size_t),
                ("sample_size", ctypes.c_size_t), ("sample_fit_error", ctypes.c_double),
                ("max_bucket_overfill", ctypes.c_double),
                ("fallback_fraction", ctypes.c_double), ("root_log", ctypes.c_int)]


class LearnedBatchedStats(ctypes.Structure):
    _fields_ = [("buckets", ctypes.c_size_t), ("sample_size", ctypes.c_size_t),
                ("fit_rejected", ctypes.c_size_t), ("max_bucket", ctypes.c_size_t),
                ("fallback_fraction", ctypes.c_double)]


# Every symbol include/compute_c.h declares, with (restype, argtypes).
_vp, _sz, _i, _cp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_char_p
SIGNATURES = {
    "clc_last_error": (_cp, []),
    "clc_last_cl_error": (_i, []),
    "clc_device_count": (_i, [ctypes.POINTER(_sz)]),
    "clc_device_get_info": (_i, [_sz, ctypes.POINTER(DeviceInfo)]),
    "clc_context_create": (_i, [_sz, ctypes.POINTER(_vp)]),
    "clc_context_create_with_cache": (_i, [_sz, _cp, ctypes.POINTER(_vp)]),
    "clc_context_create_from_stream": (_i, [_vp, ctypes.POINTER(_vp)]),
    "clc_context_release": (None, [_vp]),
    "clc_context_get_info": (_i, [_vp, ctypes.POINTER(DeviceInfo)]),
    "clc_finish": (_i, [_vp]),
    "clc_context_get_stream": (_i, [_vp, ctypes.POINTER(_vp)]),
    "clc_context_memory_stats": (_i, [_vp, ctypes.POINTER(MemoryStats)]),
    "clc_context_release_cache": (_i, [_vp]),
    "clc_context_program_stats": (_i, [_vp, ctypes.POINTER(ProgramStats)]),
    "clc_copy_if": (_i, [_vp, _vp, _vp, _sz, _i, _cp, ctypes.POINTER(_sz)]),
    "clc_predicate_supported": (_i, [_i, _cp]),
    "clc_buffer_alloc": (_i, [_vp, _sz, ctypes.POINTER(_vp)]),
    "clc_buffer_release": (None, [_vp]),
    "clc_buffer_size": (_sz, [_vp]),
    "clc_buffer_data": (_vp, [_vp]),
    "clc_buffer_write": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "clc_buffer_read": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "clc_buffer_copy": (_i, [_vp, _vp, _sz, _vp, _sz, _sz]),
    "clc_fill": (_i, [_vp, _vp, _sz, _i, _vp]),
    "clc_sequence": (_i, [_vp, _vp, _sz, _i, _vp, _vp]),
    "clc_transform": (_i, [_vp, _vp, _i, _vp, _i, _sz, _cp]),
    "clc_transform2": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i, _sz, _cp]),
    "clc_transform_supported": (_i, [_i, _i, _i, _cp]),
    "clc_context_set_profiling": (_i, [_vp, _i]),
    "clc_context_kernel_timings": (_i, [_vp, ctypes.POINTER(KernelTiming), _sz,
                                        ctypes.POINTER(_sz)]),
    "clc_context_reset_timings": (_i, [_vp]),
    "clc_reduce": (_i, [_vp, _vp, _sz, _i, _cp, _cp, _vp, _vp]),
    "clc_scan": (_i, [_vp, _vp, _vp, _sz, _i, _cp, _cp, _i]),
    "clc_scan_noalloc": (_i, [_vp, _vp, _vp, _sz, _i, _cp, _cp, _i]),
    "clc_sort": (_i, [_vp, _vp, _i, _vp, _sz, _i]),
    "clc_histogram_even": (_i, [_vp, _vp, _sz, _i, _vp, _sz, ctypes.c_double,
                                ctypes.c_double]),
    "clc_operator_supported": (_i, [_i, _cp, _cp]),
    "clc_sort_tune": (_i, [_vp, _i, _i, ctypes.POINTER(_i)]),
    "clc_segmented_sort": (_i, [_vp, _vp, _i, _vp, _sz, _vp, _sz, _i]),
    "clc_segmented_sort_tune": (_i, [_vp, _i, _i]),
    "clc_segmented_sort_get_table": (_i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_int)]),
    "clc_segmented_sort_get_digit_bits": (_i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_int)]),
    "clc_segmented_sort_graph_create": (_i, [_vp, _vp, _i, _vp, _sz, _vp, _sz, _i,
                                             ctypes.POINTER(_vp)]),
    "clc_learned_sort": (_i, [_vp, _vp, _i, _vp, _sz, _i, ctypes.POINTER(LearnedStats)]),
    "clc_learned_sort_single_scatter": (_i, [_vp, _vp, _i, _vp, _sz, _i,
                                             ctypes.POINTER(LearnedStats)]),
    "clc_learned_segmented_sort": (_i, [_vp, _vp, _i, _vp, _sz, _vp, _sz, _i,
                                        ctypes.POINTER(LearnedBatchedStats)]),
    "clc_graph_launch": (_i, [_vp]),
    "clc_graph_release": (None, [_vp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python opencl-sdk_amd/build.py` (no fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != SUCCESS:
        raise ClcError(status, lib().clc_last_error().decode(errors="replace"))


def _np_of():
    import numpy as np
    return {I8: np.int8, U8: np.uint8, I16: np.int16, U16: np.uint16, I32: np.int32,
            U32: np.uint32, I64: np.int64, U64: np.uint64, F32: np.float32, F64: np.float64}


def _enc(s):
    return None if s is None else s.encode()


def dtype_code(dtype) -> int:
    import torch
    table = {torch.int8: I8, torch.uint8: U8, torch.int16: I16,
             torch.int32: I32, torch.int64: I64, torch.float32: F32,
             torch.float64: F64}
    for name, code in (("uint16", U16), ("uint32", U32), ("uint64", U64)):
        if hasattr(torch, name):
            table[getattr(torch, name)] = code
    return table[dtype]


class Graph:
    """A captured segmented sort (clc_graph)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self):
        check(lib().clc_graph_launch(self.handle))

    def close(self):
        if self.handle:
            lib().clc_graph_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """A clc_context bound to a torch stream (default: the current one).

    With cache_dir, the context is created on device 0 with its own stream
    (clc_context_create_with_cache) and keeps run-time compiled programs and
    tuned tables in cache_dir; callers order its work against torch's with
    torch.cuda.synchronize() and finish()."""

    def __init__(self, stream=None, cache_dir=None):
        h = ctypes.c_void_p()
        if cache_dir is not None:
            self.stream = None
            check(lib().clc_context_create_with_cache(0, _enc(cache_dir), ctypes.byref(h)))
        else:
            import torch
            if stream is None:
                stream = torch.cuda.current_stream()
            self.stream = stream
            check(lib().clc_context_create_from_stream(
                ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().clc_context_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def finish(self):
        check(lib().clc_finish(self.handle))

    def program_stats(self) -> ProgramStats:
        s = ProgramStats()
        check(lib().clc_context_program_stats(self.handle, ctypes.byref(s)))
        return s

    def memory_stats(self) -> MemoryStats:
        s = MemoryStats()
        check(lib().clc_context_memory_stats(self.handle, ctypes.byref(s)))
        return s

    def set_profiling(self, on: bool):
        check(lib().clc_context_set_profiling(self.handle, 1 if on else 0))

    def kernel_timings(self) -> dict:
        count = ctypes.c_size_t()
        check(lib().clc_context_kernel_timings(self.handle, None, 0, ctypes.byref(count)))
        arr = (KernelTiming * max(count.value, 1))()
        check(lib().clc_context_kernel_timings(self.handle, arr, count.value,
                                               ctypes.byref(count)))
        return {a.name.decode(): {"launches": a.launches, "total_ms": a.total_ms,
                                  "min_ms": a.min_ms, "max_ms": a.max_ms}
                for a in arr[:count.value]}

    def reset_timings(self):
        check(lib().clc_context_reset_timings(self.handle))

    # -- algorithms (device tensors in, results in place / out) -------------

    def sort(self, keys, values=None, descending=False, key_type=None, n=None):
        kt = dtype_code(keys.dtype) if key_type is None else key_type
        n = keys.numel() if n is None else n
        check(lib().clc_sort(self.handle, ctypes.c_void_p(keys.data_ptr()), kt,
                             None if values is None else ctypes.c_void_p(values.data_ptr()),
                             n, 1 if descending else 0))

    def segmented_sort(self, keys, values, offsets, descending=False, key_type=None,
                       num_items=None, num_segments=None):
        kt = dtype_code(keys.dtype) if key_type is None else key_type
        n = keys.numel() if num_items is None else num_items
        ns = offsets.numel() - 1 if num_segments is None else num_segments
        check(lib().clc_segmented_sort(
            self.handle, ctypes.c_void_p(keys.data_ptr()), kt,
            None if values is None else ctypes.c_void_p(values.data_ptr()), n,
            ctypes.c_void_p(offsets.data_ptr()), ns, 1 if descending else 0))

    def learned_sort(self, keys, values=None, descending=False, key_type=None, n=None,
                     stats=True, single_scatter=False):
        """clc_learned_sort (clc_learned_sort_single_scatter with
        single_scatter); returns the model statistics as a dict (or None)."""
        kt = dtype_code(keys.dtype) if key_type is None else key_type
        n = keys.numel() if n is None else n
        st = LearnedStats() if stats else None
        fn = lib().clc_learned_sort_single_scatter if single_scatter else lib().clc_learned_sort
        check(fn(
            self.handle, ctypes.c_void_p(keys.data_ptr()), kt,
            None if values is None else ctypes.c_void_p(values.data_ptr()), n,
            1 if descending else 0, ctypes.byref(st) if st is not None else None))
        if st is None:
            return None
        return {name: getattr(st, name) for name, _ in LearnedStats._fields_}

    def learned_segmented_sort(self, keys, values, offsets, descending=False, key_type=None,
                               num_items=None, num_segments=None, stats=True):
        """clc_learned_segmented_sort; returns its statistics as a dict (or None)."""
        kt = dtype_code(keys.dtype) if key_type is None else key_type
        n = keys.numel() if num_items is None else num_items
        ns = offsets.numel() - 1 if num_segments is None else num_segments
        st = LearnedBatchedStats() if stats else None
        check(lib().clc_learned_segmented_sort(
            self.handle, ctypes.c_void_p(keys.data_ptr()), kt,
            None if values is None else ctypes.c_void_p(values.data_ptr()), n,
            ctypes.c_void_p(offsets.data_ptr()), ns, 1 if descending else 0,
            ctypes.byref(st) if st is not None else None))
        if st is None:
            return None
        return {name: getattr(st, name) for name, _ in LearnedBatchedStats._fields_}

    def sort_tune(self, key_type, with_values=True) -> int:
        shape = ctypes.c_int(-1)
        check(lib().clc_sort_tune(self.handle, key_type, 1 if with_values else 0,
                                  ctypes.byref(shape)))
        return shape.value

    def segmented_sort_tune(self, key_type, with_values=True):
        check(lib().clc_segmented_sort_tune(self.handle, key_type, 1 if with_values else 0))

    def segmented_sort_table(self, key_type, with_values=True):
        arr = (ctypes.c_int * 32)()
        check(lib().clc_segmented_sort_get_table(self.handle, key_type,
                                                 1 if with_values else 0, arr))
        return list(arr)

    def segmented_sort_digit_bits(self, key_type, with_values=True):
        arr = (ctypes.c_int * 32)()
        check(lib().clc_segmented_sort_get_digit_bits(self.handle, key_type,
                                                      1 if with_values else 0, arr))
        return list(arr)

    def segmented_sort_graph(self, keys, values, offsets, descending=False, key_type=None):
        kt = dtype_code(keys.dtype) if key_type is None else key_type
        h = ctypes.c_void_p()
        check(lib().clc_segmented_sort_graph_create(
            self.handle, ctypes.c_void_p(keys.data_ptr()), kt,
            None if values is None else ctypes.c_void_p(values.data_ptr()), keys.numel(),
            ctypes.c_void_p(offsets.data_ptr()), offsets.numel() - 1,
            1 if descending else 0, ctypes.byref(h)))
        return Graph(h)

    def scan(self, inp, out, op="(a + b)", identity="0", exclusive=False,
             type_code=None, n=None):
        t = dtype_code(inp.dtype) if type_code is None else type_code
        n = inp.numel() if n is None else n
        check(lib().clc_scan(self.handle, ctypes.c_void_p(inp.data_ptr()),
                             ctypes.c_void_p(out.data_ptr()), n, t, _enc(op),
                             _enc(identity), 1 if exclusive else 0))

    def reduce(self, inp, op="(a + b)", identity="0", init=0, type_code=None, n=None):
        import numpy as np
        t = dtype_code(inp.dtype) if type_code is None else type_code
        n = inp.numel() if n is None else n
        npdt = {I8: np.int8, U8: np.uint8, I16: np.int16, U16: np.uint16,
                I32: np.int32, U32: np.uint32, I64: np.int64, U64: np.uint64,
                F32: np.float32, F64: np.float64}[t]
        ini = np.array([init], dtype=npdt)
        res = np.zeros(1, dtype=npdt)
        check(lib().clc_reduce(self.handle, ctypes.c_void_p(inp.data_ptr()), n, t,
                               _enc(op), _enc(identity),
                               ini.ctypes.data_as(ctypes.c_void_p),
                               res.ctypes.data_as(ctypes.c_void_p)))
        return res[0]

    def _one(self, t, v):
        import numpy as np
        return np.array([v], dtype=_np_of()[t])

    def fill(self, out, value, type_code=None, n=None):
        t = dtype_code(out.dtype) if type_code is None else type_code
        n = out.numel() if n is None else n
        v = self._one(t, value)
        check(lib().clc_fill(self.handle, ctypes.c_void_p(out.data_ptr()), n, t,
                             v.ctypes.data_as(ctypes.c_void_p)))

    def sequence(self, out, start=0, step=1, type_code=None, n=None):
        t = dtype_code(out.dtype) if type_code is None else type_code
        n = out.numel() if n is None else n
        a, b = self._one(t, start), self._one(t, step)
        check(lib().clc_sequence(self.handle, ctypes.c_void_p(out.data_ptr()), n, t,
                                 a.ctypes.data_as(ctypes.c_void_p),
                                 b.ctypes.data_as(ctypes.c_void_p)))

    def transform(self, inp, out, expression, in_type=None, out_type=None, n=None):
        ti = dtype_code(inp.dtype) if in_type is None else in_type
        to = dtype_code(out.dtype) if out_type is None else out_type
        n = inp.numel() if n is None else n
        check(lib().clc_transform(self.handle, ctypes.c_void_p(inp.data_ptr()), ti,
                                  ctypes.c_void_p(out.data_ptr()), to, n, _enc(expression)))

    def transform2(self, in1, in2, out, expression, type1=None, type2=None, out_type=None,
                   n=None):
        t1 = dtype_code(in1.dtype) if type1 is None else type1
        t2 = dtype_code(in2.dtype) if type2 is None else type2
        to = dtype_code(out.dtype) if out_type is None else out_type
        n = in1.numel() if n is None else n
        check(lib().clc_transform2(self.handle, ctypes.c_void_p(in1.data_ptr()), t1,
                                   ctypes.c_void_p(in2.data_ptr()), t2,
                                   ctypes.c_void_p(out.data_ptr()), to, n, _enc(expression)))

    def copy_if(self, inp, out, predicate, type_code=None, n=None) -> int:
        """clc_copy_if: elements of inp satisfying predicate (an expression
        of x) to the front of out, in order; returns how many."""
        t = dtype_code(inp.dtype) if type_code is None else type_code
        n = inp.numel() if n is None else n
        count = ctypes.c_size_t(0)
        check(lib().clc_copy_if(self.handle, ctypes.c_void_p(inp.data_ptr()),
                                ctypes.c_void_p(out.data_ptr()), n, t, _enc(predicate),
                                ctypes.byref(count)))
        return count.value

    def histogram_even(self, inp, hist, lower, upper, type_code=None, n=None):
        t = dtype_code(inp.dtype) if type_code is None else type_code
        n = inp.numel() if n is None else n
        check(lib().clc_histogram_even(self.handle, ctypes.c_void_p(inp.data_ptr()),
                                       n, t, ctypes.c_void_p(hist.data_ptr()),
                                       hist.numel(), float(lower), float(upper)))
