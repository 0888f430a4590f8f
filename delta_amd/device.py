"""Device (HBM) buffers owned through the C ABI (dh_malloc / dh_free)."""
import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """A typed HBM allocation.  `ptr` is the raw device address (int)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _lib.check(_lib.lib().dh_malloc(C.byref(p), max(self.nbytes, 1)), "dh_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        _lib.check(_lib.lib().dh_memset(d.ptr, 0, d.nbytes), "dh_memset")
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        _lib.check(_lib.lib().dh_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "h2d")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _lib.check(_lib.lib().dh_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            _lib.lib().dh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
