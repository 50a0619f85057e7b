"""Device memory for the device-buffer entry points (MUTO_MEM_DEVICE,
MUTO_MAP_DEVICE_BUFFERS): a float64 array allocated with hipMalloc through
the same HIP runtime that libmuto.so links (/opt/rocm's libamdhip64), so the
pointers are valid for the library's context.  (The torch wheel carries a HIP
runtime of its own; a torch tensor's pointer is only usable when torch and
libmuto share one runtime.)"""
import atexit
import ctypes as C
import weakref

import numpy as np

_hip = None
_live = weakref.WeakSet()


@atexit.register
def _free_all():
    """Arrays still alive at exit are freed while the HIP runtime is up."""
    for a in list(_live):
        try:
            a.free()
        except Exception:
            pass


def hip():
    global _hip
    if _hip is None:
        from . import lib
        lib()  # loads libmuto.so and, with it, its HIP runtime
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _hip = h
    return _hip


H2D, D2H, D2D = 1, 2, 3  # hipMemcpyKind


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """n values of `dtype` in device memory; .ptr is the device address."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self._p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self._p), max(self.n, 1) * self.dtype.itemsize), "hipMalloc")
        _live.add(self)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._p

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype).ravel()
        assert a.size == self.n
        _ok(hip().hipMemcpy(self._p, a.ctypes.data_as(C.c_void_p), self.n * self.dtype.itemsize, H2D),
            "hipMemcpy H2D")

    def fill_bytes(self, byte):
        _ok(hip().hipMemset(self._p, int(byte), self.n * self.dtype.itemsize), "hipMemset")
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._p, self.n * self.dtype.itemsize, D2H),
            "hipMemcpy D2H")
        return out

    def free(self):
        if self._p:
            hip().hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
