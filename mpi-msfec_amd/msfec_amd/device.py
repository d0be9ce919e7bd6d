"""Device memory for the Python face without torch: a float64 buffer in HBM
allocated through the HIP runtime the library itself links, usable as an
input (``loc = MSFEC_DEVICE``) or output of every entry point that takes one.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        L.load()  # the library pulls in the runtime first: one runtime per process
        h = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        for name, args in (("hipMalloc", [ctypes.POINTER(vp), ctypes.c_size_t]),
                           ("hipFree", [vp]),
                           ("hipMemcpy", [vp, vp, ctypes.c_size_t, ctypes.c_int]),
                           ("hipGetErrorString", [ctypes.c_int])):
            getattr(h, name).argtypes = args
            getattr(h, name).restype = ctypes.c_char_p if name == "hipGetErrorString" else ctypes.c_int
        _HIP = h
    return _HIP


def _check(st, what):
    if st != 0:
        raise L.MsfecError(f"{what}: {_hip().hipGetErrorString(st).decode()}")


class DeviceBuffer:
    """``shape`` float64 values in device memory (uninitialised), or a device
    copy of ``array``."""

    def __init__(self, shape=None, array=None):
        if array is not None:
            array = np.ascontiguousarray(array, dtype=np.float64)
            shape = array.shape
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = ctypes.c_void_p()
        _check(_hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 8)), "hipMalloc")
        if array is not None:
            _check(_hip().hipMemcpy(self.ptr, ctypes.c_void_p(array.ctypes.data), self.nbytes, 1),
                   "hipMemcpy")

    def numpy(self):
        """A host copy (synchronous: waits for the work before it on the null stream)."""
        out = np.empty(self.shape)
        _check(_hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, self.nbytes, 2),
               "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
