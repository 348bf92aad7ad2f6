"""Helpers for the GPU tests: moving numpy arrays to and from HIP device
memory (through torch, which only provides memory and streams here) and a
shared clc context on the current stream."""
import sys

import numpy as np

_SIGNED_VIEW = {np.dtype(np.uint16): np.int16, np.dtype(np.uint32): np.int32,
                np.dtype(np.uint64): np.int64}


def dev(a: np.ndarray):
    import torch
    a = np.ascontiguousarray(a)
    view = _SIGNED_VIEW.get(a.dtype)
    b = a.view(view) if view is not None else a
    return torch.from_numpy(b.copy()).cuda()


def host(t, dtype) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


def say(*args):
    """Names the case on stdout before a GPU step runs."""
    print("[case]", *args, flush=True)
    sys.stdout.flush()


TYPE_OF = {np.dtype(np.int8): 0, np.dtype(np.uint8): 1, np.dtype(np.int16): 2,
           np.dtype(np.uint16): 3, np.dtype(np.int32): 4, np.dtype(np.uint32): 5,
           np.dtype(np.int64): 6, np.dtype(np.uint64): 7, np.dtype(np.float32): 8,
           np.dtype(np.float64): 9}
