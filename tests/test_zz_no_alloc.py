"""kofm_reduce_device (the reduction with no allocation) against kofm_reduce.

Each field of a context is reduced twice: by kofm_reduce over the context's
own device array, and by kofm_reduce_device over a device copy of the same
values (FP32 fields as FP32, FP64 fields as FP64).  Both run the same kernels
with the same work-group count for a given n, so the results must have the
same bits.  Two calls in a row on lattices of different size (the second
smaller) check that the shared scratch array keeps nothing that matters
between calls.

Device memory for the copies comes from the HIP runtime the library itself
runs on: the HIP functions are looked up through libkofm.so's own handle,
which searches the libraries it was linked against, so another HIP runtime
in the process (torch bundles one) is never the one called.
"""
import ctypes as C

import numpy as np
import pytest

from komegafoulingmodel_amd import channel, lib

pytestmark = pytest.mark.gpu

OPS = {"sum": lib.RED_SUM, "max": lib.RED_MAX, "min": lib.RED_MIN}


def _hip():
    """The HIP runtime functions libkofm.so itself calls."""
    h = C.CDLL(lib.load()._name)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip = hip
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.host.nbytes,
                             H2D) == 0

    def free(self):
        self.hip.hipFree(self.ptr)


def _lattice(nX, nY, seed):
    w = channel.fouled_channel(nX, nY, Fval=1e-5, seed=seed, amp=(0.0, 4.0), width=(10.0, 60.0))
    L = lib.Lattice(w.nX, w.nY, storage="f32", komega=True)
    L.set_params(w.tau0, w.Fval, w.nu, kappa_rel_tol=1e-8, kappa_max_iters=200)
    L.build_geometry(w.C0, w.BL, **w.geom_params())
    rho, ux, uy = w.initial_velocity(L.geometry("nType"))
    L.init_state(rho, ux, uy)
    L.init_komega(1e-5, 1e-3, 0.0)
    L.step(3, lib.OPTION_SAVE_MACRO_FIELDS)
    L.synchronize()
    return L


def _check(hip, L, label):
    out = DeviceArray(hip, np.full(1, np.nan))
    got = np.zeros(1)
    for field, dtype in (("rho", np.float32), ("ux", np.float32), ("k", np.float64),
                         ("omega", np.float64)):
        host = L.download(field)
        dev = DeviceArray(hip, host.astype(dtype))
        assert np.array_equal(dev.host.astype(np.float64), host)
        for name, op in OPS.items():
            ref = L.reduce(field, op)
            lib.reduce_device(dev.ptr.value, dev.host.itemsize, dev.host.size, op, out.ptr.value, 0)
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), out.ptr, 8, D2H) == 0
            print(f"{label} {field} {name}: kofm_reduce {ref!r} kofm_reduce_device {got[0]!r}",
                  flush=True)
            assert got[0] == ref, (label, field, name, got[0], ref)
        dev.free()
    out.free()


def test_reduce_device_matches_reduce_on_two_calls_the_second_smaller():
    hip = _hip()
    big = _lattice(1024, 192, seed=21)
    _check(hip, big, "1024x192")
    big.close()
    small = _lattice(256, 48, seed=22)
    _check(hip, small, "256x48")
    small.close()


def test_reduce_device_rejects_bad_arguments():
    hip = _hip()
    x = DeviceArray(hip, np.ones(16))
    out = DeviceArray(hip, np.zeros(1))
    for args in ((x.ptr.value, 8, 0, lib.RED_SUM), (x.ptr.value, 8, (1 << 40) + 1, lib.RED_SUM),
                 (x.ptr.value, 3, 16, lib.RED_SUM), (x.ptr.value, 8, 16, 3), (0, 8, 16, lib.RED_SUM)):
        with pytest.raises(lib.KofmError) as ei:
            lib.reduce_device(*args, out.ptr.value)
        assert ei.value.code == -1, args
    # a good call after the rejected ones
    lib.reduce_device(x.ptr.value, 8, 16, lib.RED_SUM, out.ptr.value)
    assert hip.hipDeviceSynchronize() == 0
    got = np.zeros(1)
    assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), out.ptr, 8, D2H) == 0
    assert got[0] == 16.0
    x.free()
    out.free()
