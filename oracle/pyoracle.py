"""ctypes wrapper for the CPU oracle (oracle/liboracle*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product package never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = 0x7FFFFFFF
SUBGROUP, HALF_COSET, HALF_ODDS = 0, 1, 2

_libs = {}

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Channel(C.Structure):
    _fields_ = [("state", C.c_uint8 * 32), ("n_draws", C.c_uint32)]


def lib(omp=False):
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    u32, u64, sz = C.c_uint32, C.c_uint64, C.c_size_t
    sig = {
        "or_m31_add": (u32, [u32, u32]),
        "or_m31_sub": (u32, [u32, u32]),
        "or_m31_mul": (u32, [u32, u32]),
        "or_m31_inv": (u32, [u32]),
        "or_cm31_mul": (None, [_u32p, _u32p, _u32p]),
        "or_qm31_mul": (None, [_u32p, _u32p, _u32p]),
        "or_qm31_inv": (None, [_u32p, _u32p]),
        "or_coset": (None, [C.c_int, u32, _u32p]),
        "or_eval_domain": (None, [u32, _u32p]),
        "or_coset_at": (None, [_u32p, u64, _u32p]),
        "or_circle_mul": (None, [_u32p, _u32p, _u32p]),
        "or_blake2s": (None, [C.c_char_p, sz, _u8p]),
        "or_blake2s_single": (None, [C.c_char_p, sz, _u8p]),
        "or_compute_twiddles": (None, [_u32p, _u32p, _u32p, _u32p, _u32p]),
        "or_batch_inverse": (None, [_u32p, _u32p, sz]),
        "or_fold_twiddles": (None, [_u32p, C.c_int, _u32p]),
        "or_ntt_layer": (None, [_u32p, _u32p, u32, u64, C.c_int]),
        "or_ntt_evaluate": (None, [_u32p, _u32p, _u32p, _u32p, u32, u64]),
        "or_ntt_interpolate": (None, [_u32p, _u32p, _u32p, _u32p, u32, u64]),
        "or_bit_reverse_m31": (None, [_u32p, u32]),
        "or_fibonacci_trace": (None, [u32, u32, u32, _u32p]),
        "or_zero_pad": (None, [_u32p, _u32p, u64, u64]),
        "or_fibonacci_quotient": (None, [_u32p, _u32p, _u32p, _u32p, _u32p, _u32p, u64]),
        "or_fold_line": (None, [C.c_void_p, _u32p, C.c_void_p, _u32p, u64]),
        "or_fold_circle_into_line": (None, [C.c_void_p, C.c_void_p, _u32p, _u32p, _u32p, u64]),
        "or_merkle_leaf_hashes": (None, [C.c_void_p, u32, u64, _u32p]),
        "or_merkle_hash_nodes": (None, [_u32p, _u32p, u64]),
        "or_merkle_root": (None, [C.c_void_p, u32, u32, _u32p]),
        "or_verify_auth_path": (C.c_int, [_u32p, _u32p, u64, _u32p, u32]),
        "or_channel_init": (None, [C.POINTER(Channel)]),
        "or_channel_mix_digest": (None, [C.POINTER(Channel), _u32p]),
        "or_channel_mix_felts": (None, [C.POINTER(Channel), _u32p, sz]),
        "or_channel_draw_felt": (None, [C.POINTER(Channel), _u32p]),
        "or_channel_draw_number": (u64, [C.POINTER(Channel), u64]),
        "or_channel_mix_u64": (None, [C.POINTER(Channel), u64]),
        "or_channel_verify_pow_nonce": (C.c_int, [C.POINTER(Channel), u32, u64]),
        "or_grind_pow": (u64, [C.POINTER(Channel), u32]),
        "or_grind_pow_stwo": (u64, [C.POINTER(Channel), u32]),
        "or_permute_hc_to_canonic_brt": (None, [_u32p, _u32p, u32]),
        "or_oods_sample": (None, [C.POINTER(Channel), C.c_void_p, _u32p, u32, u32, _u32p, _u32p]),
        "or_permute_canonic_brt_to_hc_natural": (None, [_u32p, _u32p, u32]),
        "or_pow_tz_words": (u32, [_u32p, u64, C.c_int]),
        "or_grind_pow_stwo_words": (u64, [_u32p, u32, C.c_int]),
        "or_accumulate_lift": (None, [_u32p, _u32p, u64, u32]),
        "or_lift_and_accumulate": (None, [C.c_void_p, _u32p, u32]),
        "or_prove": (C.c_int, [u32, u32, u32, u32, u32, u32,
                               C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(sz)]),
        "or_verify": (C.c_int, [C.c_char_p, sz, u32, u32, C.c_char_p, sz]),
        "or_proof_size": (sz, [u32, u32, u32, u32]),
        "or_free": (None, [C.c_void_p]),
        "or_d2_value": (u32, [u64, u32, u64]),
        "or_d2_trace": (None, [u64, u32, u32, _u32p]),
        "or_d2_coeffs": (None, [u64, u32, _u32p]),
        "or_d2_quotient": (None, [_u32p, u32, u32, u32, u64, _u32p, _u32p, _u32p, _u32p, _u32p]),
        "or_d2_prove": (C.c_int, [u64, u32, u32, u32, u64, u32, u32, u32,
                                  C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(sz)]),
        "or_d2_verify": (C.c_int, [C.c_char_p, sz, u32, u32, C.c_char_p, sz]),
        "or_d2_proof_size": (sz, [u32, u32, u32, u32]),
        "or_lde": (None, [_u32p, u32, u32, _u32p]),
        "or_ntt_evaluate_at": (None, [_u32p, u32, _u32p, C.POINTER(C.c_uint64), u32, _u32p]),
        "or_cairo_vanishing": (None, [_u32p, _u32p, _u32p, u32, u32, C.c_int]),
        "or_axpy_m31": (None, [u32, _u32p, _u32p, C.c_uint64]),
        "or_cairo_quotient": (None, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _u32p,
                                     _u32p, _u32p, _u32p, C.c_uint64, u32, C.c_uint64, C.c_int]),
        "or_cairo_quotient_chunk": (None, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                           _u32p, _u32p, _u32p, _u32p, C.c_uint64, C.c_uint64, C.c_uint64, u32]),
        "or_d2_column": (None, [u64, u32, u32, _u32p]),
        "or_batch_prove": (C.c_int, [u64, _u32p, u32, u32, u32, u32, C.POINTER(C.POINTER(C.c_uint8)),
                                     C.POINTER(sz)]),
        "or_batch_verify": (C.c_int, [C.c_char_p, sz, u32, u32, C.c_char_p, sz]),
        "or_batch_proof_size": (sz, [_u32p, u32, u32, u32]),
        "or_build_leaves_lifted": (None, [C.c_void_p, _u32p, u32, u32, _u32p]),
        "or_merkle_root_lifted": (None, [C.c_void_p, _u32p, u32, _u32p]),
        "or_stwo_twiddles": (None, [u32, _u32p, _u32p]),
        "or_stwo_ntt_evaluate": (None, [_u32p, _u32p, u32]),
        "or_stwo_ntt_interpolate": (None, [_u32p, _u32p, u32]),
        "or_stwo_eval_at_point": (None, [_u32p, u32, _u32p, _u32p, _u32p]),
        "or_eval_at_point_factors": (None, [_u32p, _u32p, u32, _u32p]),
        "or_barycentric_partials": (None, [_u32p, _u32p, u64, u32, _u32p]),
        "or_accumulate_numerators": (None, [C.c_void_p, _u32p, _u32p, _u32p, u32, u64, _u32p, _u32p, _u32p,
                                            _u32p]),
        "or_quotients_combine": (None, [_u32p, _u32p, _u32p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        _u32p, u32, _u32p, _u32p, u32, u64, _u32p, _u32p, _u32p, _u32p]),
        "or_bc_offset_row": (u32, [u32, u32, u32, C.c_int32]),
        "or_bytecode_constraint_eval": (C.c_int, [_u32p, u32, C.c_void_p, u32, u32, u32, _u32p, _u32p, _u32p,
                                                  _u32p, _u32p, _u32p]),
        "or_gkr_fix_first_variable_base": (None, [_u32p, _u32p, _u32p, u64]),
        "or_gkr_fix_first_variable_secure": (None, [_u32p, _u32p, _u32p, u64]),
        "or_gkr_gen_eq_evals_pass": (None, [_u32p, _u32p, u64]),
        "or_gkr_gen_eq_evals": (None, [_u32p, u32, _u32p, _u32p]),
        "or_gkr_next_layer": (None, [C.c_int, C.c_void_p, C.c_void_p, _u32p, _u32p, u64]),
        "or_gkr_sum_poly": (None, [C.c_int, _u32p, C.c_void_p, C.c_void_p, _u32p, u64, _u32p]),
    }
    for name_, (res, args) in sig.items():
        f = getattr(L, name_)
        f.restype = res
        f.argtypes = args
    _libs[name] = L
    return L


def u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptrs(cols):
    arr = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    return arr


# ---------------- convenience wrappers ----------------

def coset(kind, log_size):
    out = np.zeros(5, np.uint32)
    lib().or_coset(kind, log_size, out)
    return out


def eval_domain(log_size):
    out = np.zeros(5, np.uint32)
    lib().or_eval_domain(log_size, out)
    return out


def coset_at(c, i):
    out = np.zeros(2, np.uint32)
    lib().or_coset_at(u32(c), i, out)
    return int(out[0]), int(out[1])


def blake2s(data: bytes) -> bytes:
    out = np.zeros(32, np.uint8)
    lib().or_blake2s(data, len(data), out)
    return out.tobytes()


def twiddles(c):
    """Reference-layout forward twiddles of a coset: (line, circle, offsets, sizes)."""
    c = u32(c)
    log_n = int(c[4])
    n = 1 << log_n
    line = np.zeros(max(n - 1, 1), np.uint32)
    circ = np.zeros(max(n // 2, 1), np.uint32)
    offs = np.zeros(32, np.uint32)
    sizes = np.zeros(32, np.uint32)
    lib().or_compute_twiddles(c, line, circ, offs, sizes)
    nl = max(log_n - 1, 0)
    return line, circ, offs[:nl].copy(), sizes[:nl].copy()


def batch_inverse(x):
    x = u32(x)
    out = np.zeros_like(x)
    lib().or_batch_inverse(x, out, x.size)
    return out


def inverse_twiddles(c):
    line, circ, offs, sizes = twiddles(c)
    n_used = int(sizes.sum())
    iline = line.copy()
    iline[:n_used] = batch_inverse(line[:n_used])
    return iline, batch_inverse(circ), offs, sizes


def fold_twiddles(c, extract_y):
    c = u32(c)
    out = np.zeros((1 << int(c[4])) // 2, np.uint32)
    lib().or_fold_twiddles(c, 1 if extract_y else 0, out)
    return out


def evaluate(data, c):
    d = u32(data).copy()
    line, circ, offs, _ = twiddles(c)
    lib().or_ntt_evaluate(d, line, circ, u32(offs) if offs.size else np.zeros(1, np.uint32),
                          offs.size, d.size)
    return d


def interpolate(data, c):
    d = u32(data).copy()
    line, circ, offs, _ = inverse_twiddles(c)
    lib().or_ntt_interpolate(d, line, circ, u32(offs) if offs.size else np.zeros(1, np.uint32),
                             offs.size, d.size)
    return d


def ntt_layer(data, tw, layer_idx, forward):
    d = u32(data).copy()
    lib().or_ntt_layer(d, u32(tw), layer_idx, d.size, 1 if forward else 0)
    return d


def bit_reverse(data):
    d = u32(data).copy()
    lib().or_bit_reverse_m31(d, int(d.size).bit_length() - 1)
    return d


def fibonacci_trace(a, b, log_n):
    out = np.zeros(1 << log_n, np.uint32)
    lib().or_fibonacci_trace(a, b, log_n, out)
    return out


def quotient(trace, alpha):
    t = u32(trace)
    outs = [np.zeros_like(t) for _ in range(4)]
    lib().or_fibonacci_quotient(t, *outs, u32(alpha), t.size)
    return outs


def fold_line(cols, tw, alpha):
    cols = [u32(c) for c in cols]
    half = cols[0].size // 2
    outs = [np.zeros(half, np.uint32) for _ in range(4)]
    lib().or_fold_line(_ptrs(cols), u32(tw), _ptrs(outs), u32(alpha), half)
    return outs


def fold_circle_into_line(dst, src, tw, alpha, alpha_sq):
    dst = [u32(c).copy() for c in dst]
    src = [u32(c) for c in src]
    lib().or_fold_circle_into_line(_ptrs(dst), _ptrs(src), u32(tw), u32(alpha), u32(alpha_sq),
                                   dst[0].size)
    return dst


def leaf_hashes(cols):
    cols = [u32(c) for c in cols]
    out = np.zeros((cols[0].size, 8), np.uint32)
    lib().or_merkle_leaf_hashes(_ptrs(cols), len(cols), cols[0].size, out)
    return out


def hash_nodes(children):
    ch = u32(children).reshape(-1)
    out = np.zeros((ch.size // 16, 8), np.uint32)
    lib().or_merkle_hash_nodes(ch, out, ch.size // 16)
    return out


def merkle_root(cols):
    cols = [u32(c) for c in cols]
    out = np.zeros(8, np.uint32)
    lib().or_merkle_root(_ptrs(cols), len(cols), int(cols[0].size).bit_length() - 1, out)
    return out


def qm31_mul(a, b):
    out = np.zeros(4, np.uint32)
    lib().or_qm31_mul(u32(a), u32(b), out)
    return out


def prove(a, b, log_n, blowup_bits=2, n_queries=80, pow_bits=0, omp=False):
    L = lib(omp)
    p = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    rc = L.or_prove(a, b, log_n, blowup_bits, n_queries, pow_bits, C.byref(p), C.byref(n))
    if rc != 0:
        raise ValueError(f"or_prove rejected arguments (rc={rc})")
    data = C.string_at(p, n.value)
    L.or_free(p)
    return data


def verify(proof: bytes, blowup_bits=2, n_queries=80):
    err = C.create_string_buffer(256)
    ok = lib().or_verify(proof, len(proof), blowup_bits, n_queries, err, 256)
    return bool(ok), err.value.decode()


def proof_size(log_n, blowup_bits=2, n_queries=80, pow_bits=0):
    return lib().or_proof_size(log_n, blowup_bits, n_queries, pow_bits)


class PyChannel:
    def __init__(self):
        self.c = Channel()
        lib().or_channel_init(C.byref(self.c))

    @property
    def state(self):
        return bytes(self.c.state)

    def mix_digest(self, d):
        lib().or_channel_mix_digest(C.byref(self.c), u32(d))

    def mix_felts(self, felts):
        f = u32(felts).reshape(-1)
        lib().or_channel_mix_felts(C.byref(self.c), f, f.size // 4)

    def draw_felt(self):
        out = np.zeros(4, np.uint32)
        lib().or_channel_draw_felt(C.byref(self.c), out)
        return out

    def draw_number(self, bound):
        return int(lib().or_channel_draw_number(C.byref(self.c), bound))

    def mix_u64(self, v):
        lib().or_channel_mix_u64(C.byref(self.c), v)

    def verify_pow_nonce(self, bits, nonce):
        return bool(lib().or_channel_verify_pow_nonce(C.byref(self.c), bits, nonce))

    def grind(self, bits):
        return int(lib().or_grind_pow(C.byref(self.c), bits))

    def grind_stwo(self, bits):
        return int(lib().or_grind_pow_stwo(C.byref(self.c), bits))

    def oods_sample(self, cols, trace_log):
        """The OODS step (src/oods.rs): returns (points (16 words: z.x, z.y,
        z_next.x, z_next.y), values (f_k(z) for every column, then f_k(z_next)))
        and leaves the channel after the mix."""
        cs = [u32(c) for c in cols]
        logs = u32([_log2_any(c.size) for c in cs] or [0])
        pts = np.zeros(16, np.uint32)
        vals = np.zeros(max(1, 8 * len(cs)), np.uint32)
        lib().or_oods_sample(C.byref(self.c), _ptrs(cs) if cs else None, logs, len(cs), trace_log, pts, vals)
        return pts, vals[:8 * len(cs)]

    def pow_prefix_m31(self, bits):
        """The Blake2sChannelGeneric<true> prefix: pow_prefix with every word
        reduced mod P (grind_ops.rs:160-180)."""
        return self.pow_prefix(bits) % np.uint32(P)

    def pow_prefix(self, bits):
        """Blake2s(0x12345678 LE || 0^12 || state || bits LE) (src/channel.rs:140-150) as 8 words."""
        msg = np.zeros(13, np.uint32)
        msg[0] = 0x12345678
        msg[4:12] = np.frombuffer(self.state, dtype=np.uint32)
        msg[12] = bits
        return np.frombuffer(blake2s(msg.tobytes()), dtype=np.uint32).copy()


# ---- synthetic degree-2 multi-column AIR (KRAK v3) ----
D2_COLS, D2_CONS = 34, 35


def d2_trace(seed, n_cols, log_n):
    out = np.zeros(n_cols << log_n, np.uint32)
    lib().or_d2_trace(seed, n_cols, log_n, out)
    return out.reshape(n_cols, 1 << log_n)


def d2_coeffs(air_seed, n_cons):
    out = np.zeros(2 * n_cons, np.uint32)
    lib().or_d2_coeffs(air_seed, n_cons, out)
    return out


def d2_quotient(eval_cols, n_cons, air_seed, alpha):
    ev = np.ascontiguousarray(eval_cols, dtype=np.uint32)
    n_cols, N = ev.shape
    q = [np.zeros(N, np.uint32) for _ in range(4)]
    lib().or_d2_quotient(ev.reshape(-1), n_cols, N.bit_length() - 1, n_cons, air_seed, u32(alpha), *q)
    return q


def d2_prove(seed, log_n, n_cols=D2_COLS, n_cons=D2_CONS, air_seed=0, blowup_bits=2, n_queries=80, pow_bits=0,
             omp=False):
    L = lib(omp)
    p = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    rc = L.or_d2_prove(seed, log_n, n_cols, n_cons, air_seed, blowup_bits, n_queries, pow_bits, C.byref(p),
                       C.byref(n))
    if rc != 0:
        raise ValueError(f"or_d2_prove rejected arguments (rc={rc})")
    data = C.string_at(p, n.value)
    L.or_free(p)
    return data


def d2_verify(proof: bytes, blowup_bits=2, n_queries=80):
    err = C.create_string_buffer(256)
    ok = lib().or_d2_verify(proof, len(proof), blowup_bits, n_queries, err, 256)
    return bool(ok), err.value.decode()


def d2_proof_size(log_n, n_cols=D2_COLS, blowup_bits=2, n_queries=80):
    return lib().or_d2_proof_size(log_n, n_cols, blowup_bits, n_queries)


# ---- batched uneven columns ----
def evaluate_at(coeffs, dom, idx, omp=True):
    """Outputs idx of evaluate(zero-padded coeffs, dom) without the whole
    transform: O(len(coeffs)) per group of indices sharing bits >= 2."""
    c = u32(coeffs)
    log_n = c.size.bit_length() - 1
    assert c.size == 1 << log_n
    ix = np.ascontiguousarray(np.asarray(idx, dtype=np.uint64))
    out = np.zeros(ix.size, np.uint32)
    lib(omp).or_ntt_evaluate_at(c, log_n, u32(dom), ix.ctypes.data_as(C.POINTER(C.c_uint64)), ix.size, out)
    return out


def _vptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def cairo_vanishing(init, step, log_eval, log_n, inverse=True, omp=True):
    """Z_H (or 1/Z_H) of half_coset(log_n) at eval position i = the point
    init * step^bitrev(i) (cuda/cairo_constraint.cu:77-118, 990-1019)."""
    out = np.zeros(1 << log_eval, np.uint32)
    lib(omp).or_cairo_vanishing(u32(init), u32(step), out, log_eval, log_n, 1 if inverse else 0)
    return out


def axpy_m31(s, x, y):
    y = u32(y).copy()
    lib().or_axpy_m31(s, u32(x), y, y.size)
    return y


def cairo_quotient(cols, inter, alpha, vh_inv, trans, chal, blowup_step, n=None, chunk_offset=0, slab=False,
                   omp=True):
    """Quotient columns (4 x n) of cuda_cairo_quotient (slab=False: n =
    rows of the columns) or cuda_cairo_quotient_slab (columns are slabs of
    n + blowup_step rows; vh_inv / trans / output are global)."""
    cols = [u32(c) for c in cols]
    inter = [u32(c) for c in inter]
    assert len(cols) == 34 and len(inter) == 32
    if n is None:
        n = cols[0].size
    g = u32(vh_inv).size
    out = [np.zeros(g, np.uint32) for _ in range(4)]
    lib(omp).or_cairo_quotient(_vptrs(cols), _vptrs(inter), _vptrs(out), u32(alpha), u32(vh_inv), u32(trans),
                               u32(chal), n, blowup_step, chunk_offset, 1 if slab else 0)
    return out


def cairo_quotient_chunk(cols, inter12, alpha, vh_inv, trans, chal, offset, chunk_n, blowup_step, omp=True):
    cols = [u32(c) for c in cols]
    inter12 = [u32(c) for c in inter12]
    assert len(cols) == 34 and len(inter12) == 12
    out = [np.zeros(chunk_n, np.uint32) for _ in range(4)]
    lib(omp).or_cairo_quotient_chunk(_vptrs(cols), _vptrs(inter12), _vptrs(out), u32(alpha), u32(vh_inv),
                                     u32(trans), u32(chal), offset, chunk_n, cols[0].size, blowup_step)
    return out


def lde(values, blowup_bits=2):
    v = u32(values)
    log_n = v.size.bit_length() - 1
    out = np.zeros(v.size << blowup_bits, np.uint32)
    lib().or_lde(v, log_n, blowup_bits, out)
    return out


def leaves_lifted(cols, lifting_log, omp=True):
    cs = [u32(c) for c in cols]
    logs = u32([c.size.bit_length() - 1 for c in cs])
    out = np.zeros(8 << lifting_log, np.uint32)
    lib(omp).or_build_leaves_lifted(_ptrs(cs), logs, len(cs), lifting_log, out)
    return out.reshape(-1, 8)


def merkle_root_lifted(cols, omp=True):
    cs = [u32(c) for c in cols]
    logs = u32([c.size.bit_length() - 1 for c in cs])
    out = np.zeros(8, np.uint32)
    lib(omp).or_merkle_root_lifted(_ptrs(cs), logs, len(cs), out)
    return out


def d2_column(seed, col, log_n, omp=True):
    out = np.zeros(1 << log_n, np.uint32)
    lib(omp).or_d2_column(seed, col, log_n, out)
    return out


def batch_root(log_sizes, seed, blowup_bits=2, bit_reverse=True, omp=True):
    """Lifted root of the batched workload: column c = d2_column(seed, c,
    L_c), LDE at 2^blowup_bits, bit reversed, lifted tree."""
    evs = []
    for c, L in enumerate(log_sizes):
        e = np.zeros(1 << (L + blowup_bits), np.uint32)
        lib(omp).or_lde(d2_column(seed, c, L, omp), L, blowup_bits, e)
        if bit_reverse:
            lib(omp).or_bit_reverse_m31(e, L + blowup_bits)
        evs.append(e)
    return merkle_root_lifted(evs, omp)


# ---- batched-columns proof (KRAK v4) ----
def batch_prove(seed, log_sizes, blowup_bits=2, n_queries=80, pow_bits=0, omp=True):
    L = lib(omp)
    p = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    rc = L.or_batch_prove(seed, u32(log_sizes), len(log_sizes), blowup_bits, n_queries, pow_bits, C.byref(p),
                          C.byref(n))
    if rc != 0:
        raise ValueError(f"or_batch_prove rejected arguments (rc={rc})")
    data = C.string_at(p, n.value)
    L.or_free(p)
    return data


def batch_verify(proof: bytes, blowup_bits=2, n_queries=80):
    err = C.create_string_buffer(256)
    ok = lib().or_batch_verify(proof, len(proof), blowup_bits, n_queries, err, 256)
    return bool(ok), err.value.decode()


def batch_proof_size(log_sizes, blowup_bits=2, n_queries=80):
    return lib().or_batch_proof_size(u32(log_sizes), len(log_sizes), blowup_bits, n_queries)


# ---------------- stwo-basis NTT (stwo plugin PolyOps) ----------------

def permute_hc_to_canonic_brt(src):
    """Values in half_coset(log_n) natural order -> canonic bit-reversed order (src/circle.rs:192-204)."""
    s = u32(src)
    out = np.zeros_like(s)
    lib().or_permute_hc_to_canonic_brt(s, out, _log2_any(s.size))
    return out


def permute_canonic_brt_to_hc_natural(src):
    """The inverse (src/circle.rs:206-220)."""
    s = u32(src)
    out = np.zeros_like(s)
    lib().or_permute_canonic_brt_to_hc_natural(s, out, _log2_any(s.size))
    return out


def pow_tz_words(prefix, nonce, m31_out=False):
    """Trailing zeros of Blake2s(prefix || nonce)[0..16], words reduced mod P first with m31_out."""
    return int(lib().or_pow_tz_words(u32(prefix), nonce, int(m31_out)))


def grind_stwo_words(prefix, bits, m31_out=False):
    """The stwo plugin's grind over a given prefixed digest (either channel flavour)."""
    return int(lib().or_grind_pow_stwo_words(u32(prefix), bits, int(m31_out)))


def stwo_twiddles(log_n):
    """(forward, inverse) flat twiddle buffers of n/2 words for a domain of 2^log_n points."""
    assert 2 <= log_n <= 30
    fwd = np.zeros(1 << (log_n - 1), np.uint32)
    inv = np.zeros_like(fwd)
    lib().or_stwo_twiddles(log_n, fwd, inv)
    return fwd, inv


def _log2_any(n):
    log_n = n.bit_length() - 1
    assert n == 1 << log_n
    return log_n


def _log2(n):
    log_n = n.bit_length() - 1
    assert n == 1 << log_n and log_n >= 2
    return log_n


def stwo_evaluate(coeffs, tw=None, omp=False):
    d = u32(coeffs).copy()
    log_n = _log2(d.size)
    tw = stwo_twiddles(log_n)[0] if tw is None else u32(tw)
    lib(omp).or_stwo_ntt_evaluate(d, tw, log_n)
    return d


def stwo_interpolate(values, itw=None, omp=False):
    d = u32(values).copy()
    log_n = _log2(d.size)
    itw = stwo_twiddles(log_n)[1] if itw is None else u32(itw)
    lib(omp).or_stwo_ntt_interpolate(d, itw, log_n)
    return d


def stwo_eval_at_point(coeffs, zx, zy):
    """f(z) for stwo-basis coefficients at a QM31 circle point z = (zx, zy)."""
    c = u32(coeffs)
    out = np.zeros(4, np.uint32)
    lib().or_stwo_eval_at_point(c, c.size.bit_length() - 1, u32(zx), u32(zy), out)
    return tuple(int(v) for v in out)


def stwo_domain_point(log_n, j):
    """Point j of the canonic CircleDomain of 2^log_n points (half_odds(log_n - 1)
    then its conjugates)."""
    ho = coset(HALF_ODDS, log_n - 1)
    half = 1 << (log_n - 1)
    if j < half:
        return coset_at(ho, j)
    x, y = coset_at(ho, j - half)
    return x, (P - y) % P


def folding_factors(zx, zy, log_n):
    """The plugin's factors for a point (poly_ops.rs:400-410): [z.y, z.x,
    2 z.x^2 - 1, ...] (log_n entries) reversed, flattened to 4 words each."""
    maps = [tuple(int(v) for v in zy)]
    x = tuple(int(v) for v in zx)
    one = (1, 0, 0, 0)
    for _ in range(1, log_n):
        maps.append(x)
        x2 = qm31_mul(x, x)
        x = tuple((2 * int(a) - b) % P for a, b in zip(x2, one))
    maps.reverse()
    return np.array([w for m in maps for w in m] or [0], np.uint32)


def eval_at_point_factors(coeffs, factors):
    c = u32(coeffs)
    out = np.zeros(4, np.uint32)
    lib().or_eval_at_point_factors(c, u32(factors), c.size.bit_length() - 1, out)
    return tuple(int(v) for v in out)


def barycentric_partials(evals, weights, n_blocks):
    e = u32(evals)
    out = np.zeros(4 * n_blocks, np.uint32)
    lib().or_barycentric_partials(e, u32(weights), e.size, n_blocks, out)
    return out


def accumulate_numerators(cols, col_indices, b, c, n_rows):
    """out[k][row] = sum_i (c_i * cols[col_indices[i]][row] - b_i), QM31 as 4 SoA columns."""
    cs = [u32(x) for x in cols]
    out = [np.zeros(n_rows, np.uint32) for _ in range(4)]
    lib().or_accumulate_numerators(_ptrs(cs), u32(col_indices), u32(b) if len(b) else np.zeros(1, np.uint32),
                                   u32(c) if len(c) else np.zeros(1, np.uint32), len(col_indices), n_rows, *out)
    return out


def quotients_combine(sp_x, sp_y, fla, numers, acc_log_sizes, domain_xs, domain_ys, lifting_log_size):
    """numers: per accumulator, 4 SoA columns of 2^acc_log_sizes[j] words."""
    n_rows = len(domain_xs)
    ns = [[u32(col) for col in acc] for acc in numers]
    out = [np.zeros(n_rows, np.uint32) for _ in range(4)]
    lib().or_quotients_combine(u32(sp_x), u32(sp_y), u32(fla), *[_ptrs([acc[k] for acc in ns]) for k in range(4)],
                               u32(acc_log_sizes), len(ns), u32(domain_xs), u32(domain_ys), lifting_log_size,
                               n_rows, *out)
    return out


def accumulate_lift(col, curr, log_ratio):
    out = u32(col).copy()
    lib().or_accumulate_lift(out, u32(curr), out.size, log_ratio)
    return out


def lift_and_accumulate(cols):
    """One channel of the plugin's lift_and_accumulate; returns every column
    after the loop (the last is the result)."""
    cs = [u32(c).copy() for c in cols]
    logs = np.array([c.size.bit_length() - 1 for c in cs], np.uint32)
    lib().or_lift_and_accumulate(_ptrs(cs), logs, len(cs))
    return cs


def bc_offset_row(i, domain_log, eval_log, offset):
    """stwo's offset_bit_reversed_circle_domain_index (cuda/constraint_eval.cu:63-94)."""
    return int(lib().or_bc_offset_row(i, domain_log, eval_log, offset))


def bytecode_eval(words, cols, n_rows, trace_n_rows, coeff, denom_inv, acc=None):
    """cuda_bytecode_constraint_eval restated (cuda/constraint_eval.cu:97-407):
    returns the 4 accumulator columns after adding every row's weighted
    constraint sum times denom_inv[row >> trace_log]. cols: arrays or None
    (a null column reads as 0). Raises ValueError where the oracle rejects
    the program (column out of range, unknown opcode, truncated)."""
    w = u32(words)
    keep = [None if c is None else u32(c) for c in cols]
    ptrs = (C.c_void_p * max(len(keep), 1))(*[None if c is None else c.ctypes.data for c in keep])
    out = [np.zeros(n_rows, np.uint32) if acc is None else u32(acc[k]).copy() for k in range(4)]
    rc = lib().or_bytecode_constraint_eval(w, w.size, ptrs, len(keep), n_rows, trace_n_rows, u32(coeff),
                                           u32(denom_inv), *out)
    if rc != 0:
        raise ValueError("oracle rejected the bytecode")
    return out


# ---- GKR sum-check operations of the stwo plugin (cuda/gkr.cu) ----
# Secure columns are flat uint32 arrays, 4 words per QM31 element.
GKR_GRAND_PRODUCT, GKR_LOGUP_GENERIC, GKR_LOGUP_MULT, GKR_LOGUP_SINGLES = 0, 1, 2, 3


def _vp(a):
    return None if a is None else a.ctypes.data


def gkr_fix_first_variable(col, r, secure):
    """fold_mle_evals of the first variable (cuda/gkr.cu:20-74): n elements
    in (M31 or QM31), n/2 QM31 out."""
    c = u32(col)
    n = c.size // 4 if secure else c.size
    out = np.zeros(4 * (n // 2), np.uint32)
    f = lib().or_gkr_fix_first_variable_secure if secure else lib().or_gkr_fix_first_variable_base
    f(c, out, u32(r), n)
    return out


def gkr_gen_eq_evals(y, v):
    """gen_eq_evals (gkr_ops.rs:129-161): y is k QM31 values (k x 4), 2^k out."""
    yy = u32(np.asarray(y, np.uint32).reshape(-1))
    k = yy.size // 4
    out = np.zeros(4 << k, np.uint32)
    lib().or_gkr_gen_eq_evals(yy if k else np.zeros(4, np.uint32), k, u32(v), out)
    return out


def gkr_gen_eq_evals_pass(buf, y, cur):
    b = u32(buf).copy()
    lib().or_gkr_gen_eq_evals_pass(b, u32(y), cur)
    return b


def gkr_next_layer(kind, num, den, n):
    """next_layer (cuda/gkr.cu:120-212); returns (num, den) of n/2 elements
    (den None for the grand product)."""
    keep = [None if x is None else u32(x) for x in (num, den)]
    on, od = np.zeros(2 * n, np.uint32), np.zeros(2 * n, np.uint32)
    lib().or_gkr_next_layer(kind, _vp(keep[0]), _vp(keep[1]), on, od, n)
    return on, (None if kind == GKR_GRAND_PRODUCT else od)


def gkr_sum_poly(kind, eq, num, den, lam, n_terms):
    """The device sums of sum_as_poly_in_first_variable (cuda/gkr.cu:260-540):
    (eval_at_0, eval_at_2) as two 4-word arrays."""
    keep = [None if x is None else u32(x) for x in (num, den)]
    out = np.zeros(8, np.uint32)
    lib().or_gkr_sum_poly(kind, u32(eq), _vp(keep[0]), _vp(keep[1]), u32(lam if lam is not None else [0] * 4),
                          n_terms, out)
    return out[:4].copy(), out[4:].copy()
