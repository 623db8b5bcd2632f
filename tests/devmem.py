"""Device memory for GPU tests that hand device arrays to the library.

The HIP functions are looked up through libkofm.so's own handle, which
searches the libraries it was linked against, so another HIP runtime in the
process (torch bundles one) is never the one called.
"""
import ctypes as C

import numpy as np

from komegafoulingmodel_amd import lib

H2D, D2H = 1, 2


def hip():
    h = C.CDLL(lib.load()._name)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceArray:
    """A device copy of a host array; get() copies it back."""

    def __init__(self, h, host: np.ndarray):
        self.h = h
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert h.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.host.nbytes,
                           H2D) == 0

    @property
    def addr(self) -> int:
        return self.ptr.value

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.ptr)
