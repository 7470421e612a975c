"""Device buffers allocated through the HIP runtime that libdmtz.so itself runs on.

libdmtz.so links libamdhip64.so.7. Opening that soname again after the library is loaded hands
back the runtime the library already bound to, so buffers made here live in the same HIP
runtime as the library's kernels and streams, with no second runtime (such as the copy torch
bundles) in between. Calls on a context are host-synchronous, so a buffer filled here is
complete before any entry point reads it.
"""
import ctypes

import numpy as np

from . import _native as N

_HIP = None


def hip():
    """The HIP runtime libdmtz.so is bound to."""
    global _HIP
    if _HIP is None:
        N.lib()
        h = ctypes.CDLL("libamdhip64.so.7")
        P, S = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(P), S]
        h.hipFree.argtypes = [P]
        h.hipMemcpy.argtypes = [P, P, S, ctypes.c_int]
        h.hipMemset.argtypes = [P, ctypes.c_int, S]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [ctypes.c_int]
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _HIP = h
    return _HIP


H2D, D2H, D2D = 1, 2, 3


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(rc).decode()))


class DeviceArray:
    """A contiguous device array [z, y, x, ...] of a numpy dtype on one HIP device."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.device = int(device)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = ctypes.c_void_p()
        _check(hip().hipSetDevice(self.device), "hipSetDevice")
        _check(hip().hipMalloc(ctypes.byref(self._p), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._p.value

    @property
    def size(self):
        return int(np.prod(self.shape))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(hip().hipMemcpy(self._p, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self._p, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _check(hip().hipMemcpy(self._p, ctypes.c_void_p(other.ptr), self.nbytes, D2D), "hipMemcpy D2D")

    def fill_zero(self):
        _check(hip().hipMemset(self._p, 0, self.nbytes), "hipMemset")

    def free(self):
        if self._p is not None and self._p.value:
            hip().hipFree(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
