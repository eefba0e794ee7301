"""Device-buffer plumbing: numpy <-> HBM through torch tensors.

torch is used only to own HBM allocations and streams; every computation is a
call into liblzx_gpu.so.  uint32/uint64 residues travel as int32/int64 views.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_VIEW = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64, np.dtype(np.int8): np.int8,
         np.dtype(np.int32): np.int32, np.dtype(np.int64): np.int64, np.dtype(np.uint8): np.uint8}


def torch():
    import torch as _t  # noqa: WPS433 - imported lazily (slow first import)
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("no GPU visible: the lzx_amd kernels need an MI355X (gfx950)")
    return t


def to_device(arr: np.ndarray, device="cuda"):
    """Copy a numpy array to a new device tensor (bit-preserving view for uint types)."""
    t = require_gpu()
    a = np.ascontiguousarray(arr)
    view = _VIEW.get(a.dtype)
    if view is None:
        raise TypeError(f"unsupported dtype {a.dtype}")
    return t.from_numpy(a.view(view)).to(device)


def empty(shape, dtype, device="cuda"):
    t = require_gpu()
    tdt = {np.uint32: t.int32, np.uint64: t.int64, np.int8: t.int8, np.int32: t.int32}[np.dtype(dtype).type]
    return t.empty(shape, dtype=tdt, device=device)


def to_host(tensor, dtype) -> np.ndarray:
    """Copy a device tensor back and reinterpret it as `dtype`."""
    return tensor.detach().cpu().numpy().view(dtype)


def ptr(tensor) -> C.c_void_p:
    return C.c_void_p(tensor.data_ptr())


def stream_handle(stream=None) -> C.c_void_p:
    """hipStream_t of `stream` (default: torch's current stream)."""
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def synchronize():
    torch().cuda.synchronize()
