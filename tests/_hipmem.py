"""Device buffers for tests through the HIP runtime that libgiftbte_hip.so
itself links (one runtime per process; torch wheels bundle another)."""
import ctypes as C

import numpy as np

_rt = None


def rt():
    global _rt
    if _rt is None:
        from giftbte_amd import _lib
        _lib.load()  # the library's runtime is then the one dlopen returns
        _rt = C.CDLL("libamdhip64.so.7")
        _rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipDeviceSynchronize.argtypes = []
    return _rt


class DeviceArray:
    """A float64 device copy of a host array (or n NaNs)."""

    def __init__(self, host=None, n=None):
        h = np.full(n, np.nan) if host is None else np.ascontiguousarray(host, np.float64)
        self.n = h.size
        self.ptr = C.c_void_p()
        assert rt().hipMalloc(C.byref(self.ptr), max(8, h.nbytes)) == 0
        assert rt().hipMemcpy(self.ptr, h.ctypes.data, h.nbytes, 1) == 0  # host to device

    def at(self, i=0):
        return C.c_void_p(self.ptr.value + 8 * i)

    def get(self):
        h = np.empty(self.n)
        assert rt().hipDeviceSynchronize() == 0
        assert rt().hipMemcpy(h.ctypes.data, self.ptr, h.nbytes, 2) == 0  # device to host
        return h

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()
