"""Device buffers for the C ABI's device-pointer arguments (training tuples,
capture buffers).  They are allocated through the HIP runtime that
libpennant_hip.so itself links, so pointers are valid in the library."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        from . import lib
        lib()                                   # load the library (and its runtime) first
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


class DeviceBuffer:
    """A device allocation of `shape` x `dtype`, optionally filled from a host array."""

    def __init__(self, shape, dtype=np.float32, host: np.ndarray | None = None):
        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.nbytes = max(nbytes, 8)
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), self.nbytes)
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed: {rc}")
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    @classmethod
    def from_host(cls, host: np.ndarray):
        host = np.ascontiguousarray(host)
        return cls(host.shape, host.dtype, host)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.nbytes > self.nbytes:
            raise ValueError("host array larger than the buffer")
        if host.nbytes and hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy to device failed")

    def numpy(self, rows: int | None = None) -> np.ndarray:
        shape = self.shape if rows is None else (rows,) + self.shape[1:]
        out = np.empty(shape, dtype=self.dtype)
        if out.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
