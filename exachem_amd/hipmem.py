"""Minimal ctypes access to the HIP runtime the native library links
(/opt/rocm/lib/libamdhip64.so): device buffers, streams and events for the
per-task path (xt_gather_task, xt_fully_fused_ccsd_t_gpu), which takes device
pointers.  torch is not used because it bundles a second HIP runtime."""
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
        _rt.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _rt.hipEventDestroy.argtypes = [C.c_void_p]
        _rt.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _rt.hipEventSynchronize.argtypes = [C.c_void_p]
        _rt.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _rt.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _rt


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


def mem_info():
    """(free, total) device memory in bytes."""
    fr, tot = C.c_size_t(), C.c_size_t()
    _ok(rt().hipMemGetInfo(C.byref(fr), C.byref(tot)), "hipMemGetInfo")
    return fr.value, tot.value


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


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(rt().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream):
        _ok(rt().hipEventRecord(self.h, stream), "hipEventRecord")

    def elapsed_ms(self, start):
        """Milliseconds from event `start` to this one (waits for this one)."""
        _ok(rt().hipEventSynchronize(self.h), "hipEventSynchronize")
        ms = C.c_float()
        _ok(rt().hipEventElapsedTime(C.byref(ms), start.h, self.h), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        rt().hipEventDestroy(self.h)
