"""Device memory owned from Python: the counterpart of the reference's
DeviceBuffer<u32> (src/device.rs), over vx_malloc / vx_memcpy / vx_free."""
import ctypes as C

import numpy as np

from ._lib import call, load

H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """`words` uint32 of device memory (hipMalloc'd, so 256-byte aligned)."""

    def __init__(self, words):
        self.words = int(words)
        self._p = C.c_void_p()
        call("vx_malloc", C.byref(self._p), max(4 * self.words, 4))

    @classmethod
    def from_host(cls, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint32)).ravel()
        b = cls(a.size)
        if a.size:
            call("vx_memcpy", b._p, C.c_void_p(a.ctypes.data), a.nbytes, H2D)
        return b

    @classmethod
    def zeros(cls, words):
        b = cls(words)
        call("vx_memset", b._p, 0, max(4 * b.words, 4))
        return b

    def to_host(self):
        out = np.zeros(self.words, np.uint32)
        if self.words:
            call("vx_memcpy", C.c_void_p(out.ctypes.data), self._p, out.nbytes, D2H)
        return out

    def as_ptr(self):
        return self._p.value

    def __len__(self):
        return self.words

    def free(self):
        if self._p:
            load().vx_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
