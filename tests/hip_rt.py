"""Minimal ctypes access to the HIP runtime the product library links
(/opt/rocm libamdhip64.so.7), for tests that hand caller-owned device
buffers to the C ABI.  (torch bundles a second HIP runtime; loading both in
one process leaves the second without a device, so the tests do not use
torch.)"""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import pennant_amd
        pennant_amd.lib()                       # same runtime instance as the library
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device allocation holding a copy of a numpy array."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = max(host.nbytes, 8)
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), self.nbytes) == 0
        self.ptr = p.value
        if host.nbytes:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        return cls(np.zeros(shape, dtype=dtype))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None


def synchronize():
    assert hip().hipDeviceSynchronize() == 0
