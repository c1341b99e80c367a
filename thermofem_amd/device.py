"""Device buffers and a per-device execution context over the C ABI.

Plumbing for tests and the benchmark: numpy arrays are copied to device
memory owned by the library (hipMalloc through tf_alloc) and passed to the
kernel entry points as raw pointers.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

GLOBAL, PINNED, MAPPED, MANAGED = 0, 1, 2, 3


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().tf_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        check(lib().tf_ctx_synchronize(self.h))

    def close(self):
        if self.h:
            lib().tf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation with a numpy dtype and length."""

    def __init__(self, n, dtype, device=0, strategy=GLOBAL):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.strategy = strategy
        p = C.c_void_p()
        check(lib().tf_alloc(device, strategy, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_host(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, device)
        d.upload(a)
        return d

    @classmethod
    def adopt(cls, ptr, n, dtype):
        """Take ownership of a library-allocated device array."""
        d = cls.__new__(cls)
        d.n = int(n)
        d.dtype = np.dtype(dtype)
        d.strategy = GLOBAL
        d.ptr = ptr if isinstance(ptr, C.c_void_p) else C.c_void_p(ptr)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.n:
            check(lib().tf_memcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def copy_from(self, other):
        """Device-to-device copy of another array of the same size and dtype."""
        assert other.n == self.n and other.dtype == self.dtype
        if self.n:
            check(lib().tf_memcpy(self.ptr, other.ptr, self.nbytes))

    def get(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().tf_memcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            lib().tf_free(self.strategy, self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCsr:
    def __init__(self, rp, ci, v, device=0):
        self.n = len(rp) - 1
        self.nnz = len(ci)
        self.rp = DeviceArray.from_host(np.asarray(rp, np.int32), device)
        self.ci = DeviceArray.from_host(np.asarray(ci, np.int32), device)
        self.v = DeviceArray.from_host(np.asarray(v), device)

    def args(self):
        return self.rp.ptr, self.ci.ptr, self.v.ptr
