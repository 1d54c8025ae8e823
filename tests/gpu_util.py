"""Device-buffer plumbing for the GPU parity tests.

Buffers are allocated through the library's own runtime helpers (vx_malloc /
vx_memcpy, the C-ABI's counterparts of the reference's cudaMalloc/cudaMemcpy
bindings), so a test process holds exactly one HIP runtime: the one
libvortex_hip.so links."""
import ctypes as C

import numpy as np

import vortexstark_amd as vx
from vortexstark_amd._lib import call

H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, words):
        self.words = int(words)
        self._p = C.c_void_p()
        call("vx_malloc", C.byref(self._p), max(4 * self.words, 4))
        self.ptr = self._p.value

    def data_ptr(self):
        return self.ptr

    def __del__(self):
        try:
            if self._p:
                vx.load().vx_free(self._p)
                self._p = C.c_void_p()
        except Exception:
            pass


def dev(arr):
    """numpy uint32 -> device buffer (same bits)."""
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint32))
    b = DevBuf(a.size)
    if a.size:
        call("vx_memcpy", C.c_void_p(b.ptr), C.c_void_p(a.ctypes.data), a.nbytes, H2D)
    return b


def empty(n):
    b = DevBuf(n)
    call("vx_memset", C.c_void_p(b.ptr), 0, 4 * b.words)
    return b


def host(b):
    out = np.zeros(b.words, np.uint32)
    if b.words:
        call("vx_memcpy", C.c_void_p(out.ctypes.data), C.c_void_p(b.ptr), out.nbytes, D2H)
    return out


def ptr(b):
    return C.c_void_p(b.ptr)


class Pin:
    """Keeps device buffers alive for as long as the Pin is referenced.

    `ptr(dev(x))` passed straight into a call frees the buffer before the
    call runs (the temporary's last reference ends when ptr returns), so a
    kernel would read freed memory. Tests pass `pin(dev(x))` instead and keep
    `pin` until after the synchronisation that ends the kernels."""

    def __init__(self):
        self.bufs = []

    def __call__(self, b):
        self.bufs.append(b)
        return ptr(b)


def sync():
    call("vx_device_sync")


def d2h(dptr, words):
    """Copy `words` uint32 from a raw device pointer (library-owned buffers)."""
    out = np.zeros(int(words), np.uint32)
    call("vx_memcpy", C.c_void_p(out.ctypes.data), C.c_void_p(dptr), out.nbytes, D2H)
    return out
