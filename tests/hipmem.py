"""Device buffers for the GPU tests, allocated through the HIP runtime the device library is
linked to (libamdhip64.so.7 from /opt/rocm, loaded once per process).  torch ships its own HIP
runtime; two runtimes in one process do not share devices, so the tests do not hand torch
memory to the library."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


H2D, D2H = 1, 2


class DeviceBuffer:
    """n doubles of device memory."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), 8 * self.n) == 0, "hipMalloc failed"

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = cls(a.size)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, H2D) == 0

    def download(self):
        out = np.empty(self.n)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, D2H) == 0
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
