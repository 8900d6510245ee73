"""Device buffers for tests that call the library's device-pointer entry
points: hipMalloc / hipMemcpy / hipFree from the same HIP runtime that
libmgl.so links (/opt/rocm/lib/libamdhip64.so), through ctypes."""
import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("the HIP runtime library is not available")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A copy of a numpy array in device memory (freed on close / deletion)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.dtype, self.shape = a.dtype, a.shape
        self.nbytes = max(a.nbytes, 16)
        p = C.c_void_p()
        st = hip().hipMalloc(C.byref(p), self.nbytes)
        if st != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed: {st}")
        self.ptr = p.value
        if a.nbytes:
            st = hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D)
            if st != 0:
                raise RuntimeError(f"hipMemcpy failed: {st}")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            st = hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H)
            if st != 0:
                raise RuntimeError(f"hipMemcpy failed: {st}")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
