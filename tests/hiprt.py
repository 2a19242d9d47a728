"""Minimal ctypes access to the HIP runtime the native library links
(/opt/rocm/lib/libamdhip64.so): device buffers for the per-task drop-in test.
torch is not used in the GPU tests because it bundles a second HIP runtime."""
import ctypes as C

import numpy as np

_rt = None


def rt():
    global _rt
    if _rt is None:
        import exachem_amd
        exachem_amd.lib()  # the runtime our library loaded
        _rt = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        _rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _rt.hipStreamDestroy.argtypes = [C.c_void_p]
        _rt.hipStreamSynchronize.argtypes = [C.c_void_p]
    return _rt


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class DevBuf:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = max(int(nbytes), 8)
        _ok(rt().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _ok(rt().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return b

    def to_array(self, n, dtype=np.float64):
        out = np.zeros(n, dtype=dtype)
        _ok(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def zero(self):
        _ok(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(rt().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def sync(self):
        _ok(rt().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        rt().hipStreamDestroy(self.h)
