"""Device memory and streams for the tests of the device-buffer entry points:
ctypes on the HIP runtime the library itself links (the same loaded
instance), so the addresses and streams are valid for it."""
import ctypes
import os

import numpy as np

_rt = None


def rt():
    global _rt
    if _rt is None:
        path = "/opt/rocm/lib/libamdhip64.so.7"
        L = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        L.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipStreamDestroy.argtypes = [vp]
        _rt = L
    return _rt


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class DevArray:
    """A float64 device array of a given shape."""

    def __init__(self, shape, host=None):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = ctypes.c_void_p()
        _ok(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        if host is not None:
            self.upload(host)
        else:
            _ok(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        _ok(rt().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape)
        _ok(rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _ok(rt().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def sync(self):
        _ok(rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            rt().hipStreamDestroy(self.handle)
            self.handle = None
