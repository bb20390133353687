"""Device buffers for the C-ABI tests through the HIP runtime the library
itself links (libamdhip64), so one runtime serves the whole process."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """Device buffer holding a Fortran-ordered array (FP64 unless dtype says otherwise)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value

    def upload(self, a):
        a = np.asfortranarray(a, dtype=self.dtype)
        assert a.shape == self.shape
        rc = hip().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), self.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")
        return self

    def download(self):
        a = np.empty(self.shape, order="F", dtype=self.dtype)
        rc = hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), self.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return a

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()
