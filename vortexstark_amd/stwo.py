"""stwo-basis circle NTT, the transforms of the stwo Backend plugin's PolyOps.

Host mirror of src/ntt.rs:283-379 (StwoNttCache, evaluate_stwo,
interpolate_stwo) and of the plugin's batched calls
(crates/cuda-backend/src/poly_ops.rs, cuda_stwo_ntt_batch_*), over the C-ABI.
Coefficients are in stwo's basis (f = sum c_i prod_{bit b of i} m_b with
m_0 = y, m_1 = x, m_b = 2 m_{b-1}^2 - 1, src/oods.rs:133-161); values are in
bit-reversed CircleDomain order. The twiddle buffers are built on the device.

    cache = StwoNttCache(log_n)
    d = DeviceBuffer.from_host(coeffs)
    evaluate_stwo(d, cache)        # in place: coefficients -> values
    interpolate_stwo(d, cache)     # and back
"""
import ctypes as C

from ._lib import call
from .device import DeviceBuffer


def _ptr(x):
    return C.c_void_p(x if isinstance(x, int) else x.as_ptr())


def _stream(stream):
    return C.c_void_p(stream) if stream else None


class StwoNttCache:
    """Forward and inverse stwo twiddle buffers (n/2 words each) for a
    2^log_n-point domain (src/ntt.rs:287-355)."""

    def __init__(self, log_n, stream=None):
        if not 2 <= log_n <= 30:
            raise ValueError("log_n must be 2..30")
        self.log_n = int(log_n)
        half = 1 << (log_n - 1)
        self.d_fwd_twiddles = DeviceBuffer(half)
        self.d_inv_twiddles = DeviceBuffer(half)
        call("vx_stwo_twiddles", self.log_n, _ptr(self.d_fwd_twiddles), _ptr(self.d_inv_twiddles),
             _stream(stream))
        call("vx_stream_synchronize", _stream(stream))

    @property
    def n(self):
        return 1 << self.log_n


def evaluate_stwo(d_data, cache, stream=None):
    """Coefficients -> values in place (src/ntt.rs:366-371). Synchronous."""
    call("vx_stwo_ntt_evaluate", _ptr(d_data), _ptr(cache.d_fwd_twiddles), cache.n, _stream(stream))


def interpolate_stwo(d_data, cache, stream=None):
    """Values -> coefficients in place, with the 1/n scale (src/ntt.rs:374-379). Synchronous."""
    call("vx_stwo_ntt_interpolate", _ptr(d_data), _ptr(cache.d_inv_twiddles), cache.n, _stream(stream))


def _cols(columns):
    return (C.c_void_p * len(columns))(*[_ptr(c).value for c in columns])


def evaluate_stwo_batch(columns, cache, stream=None):
    """In-place evaluate of equal-size columns, one launch per pass for all of them."""
    call("vx_stwo_ntt_batch_evaluate", _cols(columns), _ptr(cache.d_fwd_twiddles), cache.n, len(columns),
         _stream(stream))


def interpolate_stwo_batch(columns, cache, stream=None):
    """In-place interpolate of equal-size columns, one launch per pass for all of them."""
    call("vx_stwo_ntt_batch_interpolate", _cols(columns), _ptr(cache.d_inv_twiddles), cache.n, len(columns),
         _stream(stream))
