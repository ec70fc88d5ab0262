"""Device buffers for GPU tests that call device-pointer entry points of the
C ABI: hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
links (the same shared object, so one runtime in the process; torch ships a
runtime of its own and is not used next to the library)."""
import ctypes as C
import os

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        from lesgo_amd import _lib
        _lib.lib()                      # load the library (and its runtime) first
        path = "/opt/rocm/lib/libamdhip64.so.7"
        if not os.path.exists(path):
            path = "libamdhip64.so"
        h = C.CDLL(path)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


H2D, D2H = 1, 2


class DeviceArray:
    """A float64 device copy of a numpy array (or an uninitialised one of a
    shape); .ptr is the device address, .get() copies it back."""

    def __init__(self, host=None, shape=None, fill=None):
        if host is not None:
            host = np.ascontiguousarray(host, dtype=np.float64)
            shape = host.shape
        elif fill is not None:
            host = np.full(shape, fill, dtype=np.float64)
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), self.nbytes)
        assert rc == 0, f"hipMalloc failed ({rc})"
        self.ptr = p
        if host is not None:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
