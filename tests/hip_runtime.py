"""Device buffers and streams for the GPU tests, through the HIP runtime that the native
library itself loads (/opt/rocm/lib/libamdhip64.so), so the tests need no second runtime in
the process.  Streams are created non-blocking, at the default priority."""
import ctypes

import numpy as np

_HIP = None
H2D, D2H = 1, 2


def hip():
    global _HIP
    if _HIP is None:
        L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreateWithFlags", "hipStreamSynchronize",
                  "hipStreamDestroy", "hipDeviceSynchronize"):
            getattr(L, f).restype = ctypes.c_int
        _HIP = L
    return _HIP


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


class DeviceArray:
    """n doubles of device memory."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(8, 8 * self.n)), "hipMalloc")

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, H2D), "hipMemcpy H2D")
        return self

    def download(self):
        out = np.empty(self.n, np.float64)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _ok(hip().hipFree(self.ptr), "hipFree")
            self.ptr = ctypes.c_void_p()


class DeviceBuffer:
    """A copy of a numpy array (any dtype) in device memory."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.dtype, self.n = a.dtype, a.size
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(8, a.nbytes)), "hipMalloc")
        if a.nbytes:
            _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n, self.dtype)
        if out.nbytes:
            _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _ok(hip().hipFree(self.ptr), "hipFree")
            self.ptr = ctypes.c_void_p()


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(hip().hipStreamCreateWithFlags(ctypes.byref(self.handle), 1), "hipStreamCreateWithFlags")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            _ok(hip().hipStreamDestroy(self.handle), "hipStreamDestroy")
            self.handle = ctypes.c_void_p()
