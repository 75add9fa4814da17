"""Device buffers for the ``_device`` entry points of libdxmc_hip.so.

The buffers come from the HIP runtime the library itself links
(libamdhip64.so.7 under /opt/rocm), so their pointers are valid in the
library's context.  PyTorch wheels bundle a HIP runtime of their own; a
process that initialises both on one GPU sees it through only one of them, so
device memory for these entry points is allocated here rather than as torch
tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def _rt() -> C.CDLL:
    global _hip
    if _hip is None:
        N.lib()   # load the library first: its runtime is the one resolved here
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
        _hip = h
    return _hip


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise N.NativeError(N.DXG_ERR_DEVICE, f"{what}: {_rt().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """``nbytes`` of device memory; ``ptr`` is its address (an int)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(_rt().hipMalloc(C.byref(p), max(int(nbytes), 1)), "hipMalloc")
        self.ptr = int(p.value)
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if offset < 0 or offset + a.nbytes > self.nbytes:
            raise ValueError("upload outside the buffer")
        _check(_rt().hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p), a.nbytes, _H2D), "hipMemcpy")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(int(count), dtype=dtype)
        if offset < 0 or offset + out.nbytes > self.nbytes:
            raise ValueError("download outside the buffer")
        _check(_rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes, _D2H), "hipMemcpy")
        return out

    def free(self) -> None:
        if self.ptr:
            _rt().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
