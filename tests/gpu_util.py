"""Device-buffer plumbing for the GPU parity tests (torch allocates, the
vx_* C-ABI computes)."""
import ctypes as C

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t

        if not t.cuda.is_available():
            raise RuntimeError("no HIP device visible")
        _torch = t
    return _torch


def dev(arr):
    """numpy uint32 -> device int32 tensor (same bits)."""
    t = torch()
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint32))
    return t.from_numpy(a.view(np.int32).copy()).to("cuda")


def empty(n):
    return torch().zeros(int(n), dtype=torch().int32, device="cuda")


def host(t):
    torch().cuda.synchronize()
    return t.cpu().numpy().view(np.uint32).copy()


def ptr(t):
    return C.c_void_p(t.data_ptr())


def sync():
    torch().cuda.synchronize()


_hip = None


def d2h(dptr, words):
    """Copy `words` uint32 from a raw device pointer (library-owned buffers)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = np.zeros(int(words), np.uint32)
    rc = _hip.hipMemcpy(out.ctypes.data, C.c_void_p(dptr), out.nbytes, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, f"hipMemcpy failed: {rc}"
    return out
