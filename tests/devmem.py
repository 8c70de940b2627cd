"""Device buffers for tests of the device-pointer entry points, through the
HIP runtime's C API (ctypes): allocate, upload, download, free."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip = L
    return _hip


class DevBuf:
    """One device allocation of `nbytes`; `ptr` is its raw address."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 1)))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def upload(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        rc = hip().hipMemcpy(ctypes.c_void_p(b.ptr), a.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(a.nbytes), 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")
        return b

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        rc = hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                             ctypes.c_size_t(out.nbytes), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
