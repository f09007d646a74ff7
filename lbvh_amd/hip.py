"""Device memory through the HIP runtime the engine library is linked to.

A process must hold one HIP runtime.  PyTorch ships its own copy of
libamdhip64 / libhsa-runtime64; if it is imported after the engine has
loaded /opt/rocm's runtime, the second runtime finds no GPU.  The engine's
Python layer therefore allocates and copies through ctypes on the runtime it
already uses (same SONAME, so dlopen returns the loaded copy), and accepts
torch tensors only when torch shares that runtime (torch imported first).
"""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from . import lib as _engine
        _engine()  # the engine pulls in its HIP runtime first
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        _hip = L
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(rc).decode()))


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    check(hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """A device buffer with a numpy dtype and shape (owned, freed on close)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def copy_from(self, src):
        """Overwrites the buffer from a host array or another DeviceArray of
        the same byte size."""
        if isinstance(src, DeviceArray):
            if src.nbytes != self.nbytes:
                raise ValueError("size mismatch")
            check(hip().hipMemcpy(self.ptr, src.ptr, self.nbytes, D2D), "hipMemcpy D2D")
            return
        a = np.ascontiguousarray(src)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def rows(self, start, stop):
        """A non-owning view of rows [start, stop) along the first axis."""
        return DeviceView(self, start, stop)

    def numel(self):
        return int(np.prod(self.shape))

    def data_ptr(self):
        return self.ptr

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceView(DeviceArray):
    """Rows [start, stop) of a DeviceArray; the parent keeps the memory."""

    def __init__(self, parent, start, stop):
        row = int(np.prod(parent.shape[1:])) if len(parent.shape) > 1 else 1
        self.parent = parent
        self.shape = (int(stop) - int(start),) + tuple(parent.shape[1:])
        self.dtype = parent.dtype
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = parent.ptr + int(start) * row * self.dtype.itemsize

    def close(self):
        self.ptr = None
