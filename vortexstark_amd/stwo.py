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


# ---- PolyOps::eval_at_point / barycentric_eval_at_point (poly_ops.rs:382-500)

P = 0x7FFFFFFF


def qm31_mul(x, y):
    """QM31 product, words [a.re, a.im, b.re, b.im] with u^2 = 2 + i."""
    def cm(a, b):
        return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)
    a, b, c, d = (x[0], x[1]), (x[2], x[3]), (y[0], y[1]), (y[2], y[3])
    ac, bd, ad, bc = cm(a, c), cm(b, d), cm(a, d), cm(b, c)
    bdu = cm(bd, (2, 1))
    return ((ac[0] + bdu[0]) % P, (ac[1] + bdu[1]) % P, (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P)


def folding_factors(zx, zy, log_n):
    """[z.y, z.x, 2 z.x^2 - 1, ...] (log_n entries) reversed, as the plugin
    builds them (poly_ops.rs:400-410); 4 words per factor."""
    maps = [tuple(int(v) % P for v in zy)]
    x = tuple(int(v) % P for v in zx)
    for _ in range(1, log_n):
        maps.append(x)
        x2 = qm31_mul(x, x)
        x = ((2 * x2[0] - 1) % P, 2 * x2[1] % P, 2 * x2[2] % P, 2 * x2[3] % P)
    maps.reverse()
    return [w for m in maps for w in m]


def eval_at_point(d_coeffs, n, zx, zy, stream=None):
    """f(z) of n stwo-basis coefficients on the device at the QM31 point
    z = (zx, zy) (4-word tuples); returns the QM31 as a 4-tuple."""
    log_n = n.bit_length() - 1
    if n <= 0 or n != 1 << log_n:
        raise ValueError("n must be a power of two")
    f = folding_factors(zx, zy, log_n) or [0, 0, 0, 0]
    d_f = DeviceBuffer.from_host(f)
    words = 4 if n < (1 << 14) else max(80, n // 1024)
    s1, s2 = DeviceBuffer(words), DeviceBuffer(words)
    res = (C.c_uint32 * 4)()
    call("vx_eval_at_point", _ptr(d_coeffs), _ptr(d_f), res, n, _ptr(s1), _ptr(s2), _stream(stream))
    return tuple(res)


def barycentric_eval(d_evals, d_weights, n, stream=None):
    """sum_i evals[i] * weights[i] (weights QM31, 4 words each): at most 256
    blocks of partial sums on the device, added here (poly_ops.rs:445-500)."""
    if n == 0:
        return (0, 0, 0, 0)
    n_blocks = min((n + 255) // 256, 256)
    d_out = DeviceBuffer(4 * n_blocks)
    call("vx_barycentric_eval", _ptr(d_evals), _ptr(d_weights), n, _ptr(d_out), n_blocks, _stream(stream))
    parts = d_out.to_host().reshape(n_blocks, 4).astype(object)
    return tuple(int(v) % P for v in parts.sum(axis=0))


# ---- QuotientOps, GrindOps, column and accumulation ops

def _dev_ptr_array(bufs):
    """A device array holding the device addresses of `bufs` (kept by the caller)."""
    import numpy as np
    return DeviceBuffer.from_host(np.array([b.as_ptr() for b in bufs], np.uint64).view(np.uint32))


def eval_at_point_batch(columns, log_sizes, zx, zy, stream=None):
    """f_k(z) for every column (the OODS step): one call for all columns;
    returns a list of QM31 4-tuples."""
    import numpy as np
    max_log = max(log_sizes)
    d_f = DeviceBuffer.from_host(folding_factors(zx, zy, max_log) or [0, 0, 0, 0])
    d_res = DeviceBuffer(4 * len(columns))
    ptrs = (C.c_void_p * len(columns))(*[b.as_ptr() for b in columns])
    logs = (C.c_uint32 * len(columns))(*log_sizes)
    call("vx_eval_at_point_batch", ptrs, logs, len(columns), _ptr(d_f), max_log, _ptr(d_res), _stream(stream))
    return [tuple(int(v) for v in r) for r in d_res.to_host().reshape(-1, 4)]


def accumulate_numerators(columns, col_indices, b, c, n_rows, stream=None):
    """QuotientOps::accumulate_numerators (quotient_ops.rs:120-208): the 4
    SoA output columns of sum_i (c_i * columns[col_indices[i]][row] - b_i);
    b, c are flat QM31 word lists. Returns 4 DeviceBuffers."""
    import numpy as np
    d_cols = _dev_ptr_array(columns)
    d_idx = DeviceBuffer.from_host(np.array(col_indices or [0], np.uint32))
    d_b = DeviceBuffer.from_host(np.array(b or [0, 0, 0, 0], np.uint32))
    d_c = DeviceBuffer.from_host(np.array(c or [0, 0, 0, 0], np.uint32))
    out = [DeviceBuffer(n_rows) for _ in range(4)]
    call("vx_accumulate_numerators", _ptr(d_cols), _ptr(d_idx), _ptr(d_b), _ptr(d_c), len(col_indices), n_rows,
         *[_ptr(o) for o in out], _stream(stream))
    call("vx_stream_synchronize", _stream(stream))  # the staging buffers above are freed on return
    return out


def grind(prefixed_digest_words, pow_bits, stream=None, m31_output=False):
    """GrindOps (grind_ops.rs:21-135): the smallest valid nonce (hi << 32) | low,
    low < 2^20, for the 8-word prefixed digest. m31_output selects the
    Blake2sChannelGeneric<true> flavour (grind_ops.rs:154-245): the hash words
    are reduced mod P before the trailing-zero count, and the caller passes
    the prefixed digest already reduced the same way."""
    import numpy as np
    d_p = DeviceBuffer.from_host(np.asarray(prefixed_digest_words, np.uint32))
    d_r = DeviceBuffer(2)
    call("vx_grind_pow_stwo_m31" if m31_output else "vx_grind_pow_stwo", _ptr(d_p), pow_bits, _ptr(d_r),
         _stream(stream))
    call("vx_stream_synchronize", _stream(stream))
    r = d_r.to_host()
    return int(r[0]) | (int(r[1]) << 32)


def lift_and_accumulate(channels, log_sizes, stream=None):
    """AccumulationOps::lift_and_accumulate (accumulation_ops.rs:44-85):
    channels[c][k] is coordinate c of column k (ascending sizes); the last
    column of every channel receives the lifted sum, in place."""
    flat = [b.as_ptr() for ch in channels for b in ch]
    call("vx_lift_and_accumulate", (C.c_void_p * len(flat))(*flat), (C.c_uint32 * len(log_sizes))(*log_sizes),
         len(log_sizes), _stream(stream))


def permute_hc_to_canonic_brt(d_src, log_n, stream=None):
    """Coset::permute_hc_natural_to_canonic_brt on the device (src/circle.rs:
    192-204, cuda_permute_hc_to_canonic_brt): values in half_coset(log_n)
    natural order into the canonic bit-reversed order the stwo NTT takes.
    Returns a new DeviceBuffer."""
    out = DeviceBuffer(1 << log_n)
    call("vx_permute_hc_to_canonic_brt", _ptr(d_src), _ptr(out), 1 << log_n, log_n, _stream(stream))
    return out


def permute_canonic_brt_to_hc_natural(d_src, log_n, stream=None):
    """The inverse (src/circle.rs:206-220, cuda_permute_canonic_brt_to_hc_natural).
    Returns a new DeviceBuffer."""
    out = DeviceBuffer(1 << log_n)
    call("vx_permute_canonic_brt_to_hc_natural", _ptr(d_src), _ptr(out), 1 << log_n, log_n, _stream(stream))
    return out


def oods_sample(d_channel, columns, log_sizes, trace_log_n, stream=None):
    """The OODS step of src/oods.rs on the device: d_channel is a 9-word
    DeviceBuffer (channel state as 8 LE words, then n_draws), updated in
    place; columns are stwo-basis coefficient DeviceBuffers. Returns
    (points, values): z.x, z.y, z_next.x, z_next.y (16 words) and every
    f_k(z) then every f_k(z_next) (4 words each)."""
    n = len(columns)
    d_pts = DeviceBuffer(16)
    d_vals = DeviceBuffer(max(1, 8 * n))
    ptrs = (C.c_void_p * max(1, n))(*[c.as_ptr() for c in columns])
    logs = (C.c_uint32 * max(1, n))(*log_sizes)
    call("vx_oods_sample", _ptr(d_channel), ptrs, logs, n, trace_log_n, _ptr(d_pts), _ptr(d_vals), _stream(stream))
    return d_pts.to_host(), d_vals.to_host()[:8 * n]
