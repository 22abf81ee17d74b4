"""Device (HBM) buffers for the entry points that take device pointers.

A thin ctypes view of the HIP runtime that libsp_hip.so itself links
(libamdhip64.so.7), so that buffers and library share one runtime instance;
no other GPU framework is loaded into the process."""
import ctypes as C

import numpy as np

from . import _abi

_HOST_TO_DEVICE, _DEVICE_TO_HOST = 1, 2
_hip = None


def _rt():
    global _hip
    if _hip is None:
        _abi.load()   # maps the library's own HIP runtime into the process
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {_rt().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A C-contiguous array in device memory; `ptr` is its device address."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(_rt().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _check(_rt().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, _HOST_TO_DEVICE), "hipMemcpy")
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _check(_rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _DEVICE_TO_HOST), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            _check(_rt().hipFree(self.ptr), "hipFree")
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
