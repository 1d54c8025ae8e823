"""The stwo plugin's QuotientOps on the GPU through the C-ABI, bit-exact
against the oracle (cuda/quotient.cu restated in oracle/vx_oracle.c, pinned
by tests/test_oracle_stwo.py)."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import vortexstark_amd as vx
from gpu_util import Pin, dev, empty, host, ptr, sync
from vortexstark_amd._lib import VortexError, call

pytestmark = pytest.mark.gpu
P = po.P
NULL = None
NONE = 0xFFFFFFFF


def rnd(n, seed):
    return np.random.default_rng(seed).integers(0, P, int(n), dtype=np.uint32)


def dev_ptrs(bufs):
    """device array of the device addresses of `bufs`"""
    return dev(np.array([b.ptr for b in bufs], np.uint64).view(np.uint32))


@pytest.fixture(scope="module", autouse=True)
def _lib_loaded():
    vx.load()
    print("stwo quotient-op tests: library loaded", flush=True)


@pytest.mark.parametrize("n_rows,idx", [(1, [0]), (100, [2, 0, 2]), (4096, []), (4096, [5, 1, 3, 3, 0]),
                                        (1 << 20, [0, 1, 2, 3, 4, 5, 0, 1, 2])])
def test_accumulate_numerators(n_rows, idx):
    print(f"accumulate_numerators rows={n_rows} cols={len(idx)}", flush=True)
    cols = [rnd(n_rows, 10 + c) for c in range(6)]
    bufs = [dev(c) for c in cols]
    b, c = rnd(4 * len(idx), 1), rnd(4 * len(idx), 2)
    outs = [empty(n_rows) for _ in range(4)]
    pin = Pin()
    call("vx_accumulate_numerators", pin(dev_ptrs(bufs)), pin(dev(np.array(idx or [0], np.uint32))),
         pin(dev(b if len(idx) else np.zeros(4, np.uint32))), pin(dev(c if len(idx) else np.zeros(4, np.uint32))),
         len(idx), n_rows, *[ptr(o) for o in outs], NULL)
    sync()
    want = po.accumulate_numerators(cols, idx, b, c, n_rows)
    for k in range(4):
        assert np.array_equal(host(outs[k]), want[k]), k


@pytest.mark.parametrize("n_rows", [1001, 4096])
def test_accumulate_numerators_unaligned_column(n_rows):
    # column 1 starts off 16-byte alignment (word loads inside the 4-row
    # path); 1001 rows (not a multiple of 4) take the strided-row kernel
    cols = [rnd(n_rows, 60 + c) for c in range(3)]
    bufs = [dev(cols[0]), dev(np.concatenate([[5], cols[1]]).astype(np.uint32)), dev(cols[2])]
    addrs = np.array([bufs[0].ptr, bufs[1].ptr + 4, bufs[2].ptr], np.uint64).view(np.uint32)
    idx = [1, 0, 2, 1]
    b, c = rnd(16, 61), rnd(16, 62)
    outs = [empty(n_rows) for _ in range(4)]
    pin = Pin()
    call("vx_accumulate_numerators", pin(dev(addrs)), pin(dev(np.array(idx, np.uint32))), pin(dev(b)), pin(dev(c)),
         len(idx), n_rows, *[ptr(o) for o in outs], NULL)
    sync()
    want = po.accumulate_numerators(cols, idx, b, c, n_rows)
    for k in range(4):
        assert np.array_equal(host(outs[k]), want[k]), k


@pytest.mark.parametrize("n_rows", [256, 5000, 1 << 18])
def test_accumulate_numerators_dual(n_rows):
    cols = [rnd(n_rows, 20 + c) for c in range(5)]
    bufs = [dev(c) for c in cols]
    idx_z = [4, 0, 1, 3, 2, 0]
    zn_idx = [1, NONE, 0, NONE, 2, 3]  # z_next coefficients for 4 of the 6 columns
    b_z, c_z = rnd(24, 3), rnd(24, 4)
    b_zn, c_zn = rnd(16, 5), rnd(16, 6)
    oz = [empty(n_rows) for _ in range(4)]
    on = [empty(n_rows) for _ in range(4)]
    pin = Pin()
    call("vx_accumulate_numerators_dual", pin(dev_ptrs(bufs)), pin(dev(np.array(idx_z, np.uint32))),
         pin(dev(np.array(zn_idx, np.uint32))), pin(dev(b_z)), pin(dev(c_z)), pin(dev(b_zn)), pin(dev(c_zn)),
         len(idx_z), n_rows, *[ptr(o) for o in oz + on], NULL)
    sync()
    want_z = po.accumulate_numerators(cols, idx_z, b_z, c_z, n_rows)
    sel = [i for i, z in enumerate(zn_idx) if z != NONE]
    bn = np.concatenate([b_zn[4 * zn_idx[i]:4 * zn_idx[i] + 4] for i in sel])
    cn = np.concatenate([c_zn[4 * zn_idx[i]:4 * zn_idx[i] + 4] for i in sel])
    want_n = po.accumulate_numerators(cols, [idx_z[i] for i in sel], bn, cn, n_rows)
    for k in range(4):
        assert np.array_equal(host(oz[k]), want_z[k])
        assert np.array_equal(host(on[k]), want_n[k])


def run_combine(sx, sy, fla, numers, logs, xs, ys, L):
    n_rows = len(xs)
    nb = [[dev(col) for col in acc] for acc in numers]
    outs = [empty(n_rows) for _ in range(4)]
    pin = Pin()
    call("vx_compute_quotients_combine", pin(dev(sx)), pin(dev(sy)), pin(dev(fla)),
         *[pin(dev_ptrs([acc[k] for acc in nb])) for k in range(4)], pin(dev(np.array(logs, np.uint32))),
         len(numers), pin(dev(xs)), pin(dev(ys)), L, n_rows, *[ptr(o) for o in outs], NULL)
    sync()
    return [host(o) for o in outs]


@pytest.mark.parametrize("L,logs", [(1, [1]), (5, [5, 4]), (12, [12, 11, 9]), (20, [20, 19]), (21, [21, 20, 17])])
def test_compute_quotients_combine(L, logs):
    print(f"quotients_combine L={L} accs={len(logs)}", flush=True)
    rng = np.random.default_rng(L)
    n = 1 << L
    sx, sy, fla = (rng.integers(0, P, 4 * len(logs), dtype=np.uint32) for _ in range(3))
    numers = [[rng.integers(0, P, 1 << g, dtype=np.uint32) for _ in range(4)] for g in logs]
    xs, ys = rnd(n, 7), rnd(n, 8)
    got = run_combine(sx, sy, fla, numers, logs, xs, ys, L)
    want = po.quotients_combine(sx, sy, fla, numers, logs, xs, ys, L)
    for k in range(4):
        assert np.array_equal(got[k], want[k]), k


def test_combine_zero_denominators_and_partial_rows():
    # Im p.x = 0, Im p.y = 1, Re p.x = (x of row 5, 0): the denominator
    # Re p.x - d.x vanishes on the rows sharing that x, and only there
    L, n = 11, 2000  # n_rows below 2^L: a partial last workgroup
    xs, ys = rnd(n, 9), rnd(n, 10)
    xs[[17, 1000, 1999]] = xs[5]
    sx = np.array([xs[5], 0, 0, 0, 11, 12, 13, 14], np.uint32)
    sy = np.array([3, 4, 1, 0, 15, 16, 17, 18], np.uint32)
    fla = rnd(8, 11)
    numers = [[rnd(1 << L, 20 + k) for k in range(4)], [rnd(1 << (L - 2), 30 + k) for k in range(4)]]
    got = run_combine(sx, sy, fla, numers, [L, L - 2], xs, ys, L)
    want = po.quotients_combine(sx, sy, fla, numers, [L, L - 2], xs, ys, L)
    for k in range(4):
        assert np.array_equal(got[k], want[k]), k


def test_combine_single_row_accumulator():
    # log size 0 lies outside the lifting formula; the kernel clamps the
    # index to the one row, which reads like a 2-row accumulator [v, v]
    L, n = 11, 1 << 11
    rng = np.random.default_rng(5)
    sx, sy, fla = (rng.integers(0, P, 8, dtype=np.uint32) for _ in range(3))
    xs, ys = rnd(n, 12), rnd(n, 13)
    big = [rnd(1 << L, 40 + k) for k in range(4)]
    one = [rnd(1, 50 + k) for k in range(4)]
    got = run_combine(sx, sy, fla, [big, one], [L, 0], xs, ys, L)
    want = po.quotients_combine(sx, sy, fla, [big, [np.repeat(c, 2) for c in one]], [L, 1], xs, ys, L)
    for k in range(4):
        assert np.array_equal(got[k], want[k]), k


def test_quotient_ops_reject_bad_arguments():
    pin = Pin()
    small, rows = empty(8), empty(64)
    with pytest.raises(VortexError):  # more rows than the lifting domain
        call("vx_compute_quotients_combine", *[pin(small)] * 8, 1, pin(rows), pin(rows), 5, 64, *[pin(rows)] * 4,
             NULL)
    with pytest.raises(VortexError):
        call("vx_accumulate_numerators", NULL, NULL, NULL, NULL, 2, 16, *[pin(rows)] * 4, NULL)


# ---- GrindOps (grind_ops.rs:21-135): the (hi << 32) | low nonce family

@pytest.mark.parametrize("seed,bits", [(0, 1), (0, 8), (3, 16), (0, 21), (1, 21), (2, 21)])
def test_grind_pow_stwo(seed, bits):
    print(f"grind stwo seed={seed} bits={bits}", flush=True)
    ch = po.PyChannel()
    ch.mix_digest([seed, 1, 2, 3, 4, 5, 6, 7])
    dp, res = dev(ch.pow_prefix(bits)), empty(2)
    call("vx_grind_pow_stwo", ptr(dp), bits, ptr(res), NULL)
    sync()
    got = host(res)
    nonce = int(got[0]) | (int(got[1]) << 32)
    assert nonce == ch.grind_stwo(bits)
    assert ch.verify_pow_nonce(bits, nonce)
    # the plugin's own loop: reference-style batches of 2^20 at (hi << 32)
    for hi in range(nonce >> 32):
        r = dev(np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32))
        call("vx_grind_pow", ptr(dp), ptr(r), bits, hi << 32, 1 << 20, NULL)
        sync()
        assert host(r).tolist() == [0xFFFFFFFF, 0xFFFFFFFF]
    r = dev(np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32))
    call("vx_grind_pow", ptr(dp), ptr(r), bits, (nonce >> 32) << 32, 1 << 20, NULL)
    sync()
    g = host(r)
    assert int(g[0]) | (int(g[1]) << 32) == nonce


def test_grind_pow_stwo_zero_bits():
    res = dev(np.array([7, 7], np.uint32))
    dp = dev(np.zeros(8, np.uint32))
    call("vx_grind_pow_stwo", ptr(dp), 0, ptr(res), NULL)
    assert host(res).tolist() == [0, 0]
    res = dev(np.array([7, 7], np.uint32))
    call("vx_grind_pow_stwo_m31", ptr(dp), 0, ptr(res), NULL)
    assert host(res).tolist() == [0, 0]


# GrindOps for Blake2sChannelGeneric<true> (grind_ops.rs:154-245,
# cuda/grind_m31_output.cu): hash words reduced mod P before the count
@pytest.mark.parametrize("seed,bits", [(0, 1), (0, 8), (3, 16), (0, 21), (1, 21), (4, 22)])
def test_grind_pow_stwo_m31(seed, bits):
    print(f"grind stwo m31 seed={seed} bits={bits}", flush=True)
    ch = po.PyChannel()
    ch.mix_digest([seed, 9, 2, 3, 4, 5, 6, 7])
    pre = ch.pow_prefix_m31(bits)
    dp, res = dev(pre), empty(2)
    call("vx_grind_pow_stwo_m31", ptr(dp), bits, ptr(res), NULL)
    sync()
    got = host(res)
    nonce = int(got[0]) | (int(got[1]) << 32)
    assert nonce == po.grind_stwo_words(pre, bits, True)
    assert po.pow_tz_words(pre, nonce, True) >= bits
    # the plugin's loop over cuda_grind_pow_m31_output batches of 2^20
    for hi in range((nonce >> 32) + 1):
        r = dev(np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32))
        call("vx_grind_pow_m31_output", ptr(dp), ptr(r), bits, hi << 32, 1 << 20, NULL)
        sync()
        g = host(r)
        want = nonce if hi == nonce >> 32 else (1 << 64) - 1
        assert int(g[0]) | (int(g[1]) << 32) == want
    # the same prefix under the raw test finds the raw flavour's nonce
    call("vx_grind_pow_stwo", ptr(dp), bits, ptr(res), NULL)
    sync()
    got = host(res)
    assert int(got[0]) | (int(got[1]) << 32) == po.grind_stwo_words(pre, bits, False)


# ---- column ops (column.rs, accumulation_ops.rs)

@pytest.mark.parametrize("n", [1, 3, 4, 1000, 1 << 20])
def test_m31_add_mul(n):
    a, b = rnd(n, 70), rnd(n, 71)
    a[:1] = P - 1
    b[:1] = P - 1
    da, db, do = dev(a), dev(b), empty(n)
    call("vx_m31_add", ptr(da), ptr(db), ptr(do), n, NULL)
    sync()
    assert np.array_equal(host(do), (a.astype(np.uint64) + b) % P)
    call("vx_m31_mul", ptr(da), ptr(db), ptr(do), n, NULL)
    sync()
    assert np.array_equal(host(do), (a.astype(np.uint64) * b) % P)
    if n >= 8:  # unaligned operands take the word path
        d1 = dev(np.concatenate([[0], a]).astype(np.uint32))
        call("vx_m31_add", C.c_void_p(d1.ptr + 4), ptr(db), ptr(do), n, NULL)
        sync()
        assert np.array_equal(host(do), (a.astype(np.uint64) + b) % P)


@pytest.mark.parametrize("col_log,lr", [(1, 0), (5, 2), (20, 3), (21, 0)])
def test_accumulate_lift(col_log, lr):
    col, curr = rnd(1 << col_log, 72), rnd(1 << (col_log - lr), 73)
    dc, du = dev(col), dev(curr)
    call("vx_accumulate_lift", ptr(dc), ptr(du), 1 << col_log, lr, NULL)
    sync()
    assert np.array_equal(host(dc), po.accumulate_lift(col, curr, lr))


@pytest.mark.parametrize("logs", [[4], [1, 3, 3, 5, 8], [10, 12, 16, 20], [14] * 5 + [22]])
def test_lift_and_accumulate(logs):
    print(f"lift_and_accumulate logs={logs}", flush=True)
    chans = [[rnd(1 << L, 100 * c + k) for k, L in enumerate(logs)] for c in range(4)]
    bufs = [[dev(x) for x in ch] for ch in chans]
    arr = (C.c_void_p * (4 * len(logs)))(*[b.ptr for ch in bufs for b in ch])
    call("vx_lift_and_accumulate", arr, (C.c_uint32 * len(logs))(*logs), len(logs), NULL)
    for c in range(4):
        want = po.lift_and_accumulate(chans[c])
        assert np.array_equal(host(bufs[c][-1]), want[-1]), c
        for k in range(len(logs) - 1):  # smaller columns are only read
            assert np.array_equal(host(bufs[c][k]), chans[c][k])
    with pytest.raises(VortexError):  # sizes must ascend
        call("vx_lift_and_accumulate", arr, (C.c_uint32 * 2)(5, 4), 2, NULL)


@pytest.mark.parametrize("words", [1, 8])
def test_gather(words):
    n_src, n = 5000, 3000
    src = rnd(n_src * words, 74)
    idx = np.random.default_rng(75).integers(0, n_src, n, dtype=np.uint32)
    fn = "vx_gather_u32" if words == 1 else "vx_gather_u256"
    ds, di, dd = dev(src), dev(idx), empty(n * words)
    call(fn, ptr(ds), ptr(di), ptr(dd), n, NULL)
    sync()
    assert np.array_equal(host(dd).reshape(n, words), src.reshape(n_src, words)[idx])
    # checked form: indices at or beyond src_len (in words) leave dst as it was
    idx2 = idx.copy()
    idx2[::7] = n_src + 5
    idx2[1] = n_src - 1
    before = rnd(n * words, 76)
    di2, dd2 = dev(idx2), dev(before)
    call(fn + "_checked", ptr(ds), ptr(di2), ptr(dd2), n, n_src * words, NULL)
    sync()
    got = host(dd2).reshape(n, words)
    ok = idx2 < n_src
    assert np.array_equal(got[ok], src.reshape(n_src, words)[idx2[ok]])
    assert np.array_equal(got[~ok], before.reshape(n, words)[~ok])


@pytest.mark.parametrize("n", [16, 4096, 1 << 20])
def test_pack_leaves(n):
    ins = [rnd(n, 80 + c) for c in range(4)]
    din = [dev(x) for x in ins]
    dout = [empty(n // 16) for _ in range(64)]
    pi, po_ = dev_ptrs(din), dev_ptrs(dout)
    call("vx_pack_leaves", ptr(pi), ptr(po_), n, NULL)
    sync()
    for c in range(4):
        for m in range(16):
            # out[16 c + m][r] = in[c][16 r + m]
            assert np.array_equal(host(dout[16 * c + m]), ins[c][m::16]), (c, m)
    with pytest.raises(VortexError):
        call("vx_pack_leaves", ptr(pi), ptr(po_), n + 8, NULL)


def test_python_mirror_of_the_plugin_ops():
    from vortexstark_amd import stwo
    from vortexstark_amd.device import DeviceBuffer
    n = 4096
    cols = [rnd(n, 200 + i) for i in range(3)]
    dcols = [DeviceBuffer.from_host(c) for c in cols]
    b, c = rnd(8, 210).tolist(), rnd(8, 211).tolist()
    out = stwo.accumulate_numerators(dcols, [2, 0], b, c, n)
    want = po.accumulate_numerators(cols, [2, 0], b, c, n)
    assert all(np.array_equal(o.to_host(), w) for o, w in zip(out, want))
    ch = po.PyChannel()
    ch.mix_digest([7, 7, 7, 7, 7, 7, 7, 7])
    assert stwo.grind(ch.pow_prefix(12), 12) == ch.grind_stwo(12)
    assert stwo.grind(ch.pow_prefix_m31(12), 12, m31_output=True) == po.grind_stwo_words(ch.pow_prefix_m31(12), 12, True)
    chans = [[DeviceBuffer.from_host(rnd(1 << L, 300 + 10 * k + L)) for L in (3, 6)] for k in range(4)]
    host_chans = [[b_.to_host() for b_ in ch_] for ch_ in chans]
    stwo.lift_and_accumulate(chans, [3, 6])
    for k in range(4):
        assert np.array_equal(chans[k][1].to_host(), po.lift_and_accumulate(host_chans[k])[1])
    zx, zy = (1, 2, 3, 4), (5, 6, 7, 8)
    vals = stwo.eval_at_point_batch(dcols, [12, 12, 12], zx, zy)
    assert vals == [po.stwo_eval_at_point(x, zx, zy) for x in cols]


# ---- half-coset natural <-> canonic bit-reversed order (cuda/permute.cu,
# src/circle.rs:192-220): scatter path up to log_n 10, tiles from 11

@pytest.mark.parametrize("log_n", [0, 1, 2, 5, 10, 11, 12, 13, 16, 20, 24])
def test_permute_hc_brt(log_n):
    print(f"permute log_n={log_n}", flush=True)
    n = 1 << log_n
    src = rnd(n, 90 + log_n)
    ds, dd, dk = dev(src), empty(n), empty(n)
    call("vx_permute_hc_to_canonic_brt", ptr(ds), ptr(dd), n, log_n, NULL)
    call("vx_permute_canonic_brt_to_hc_natural", ptr(dd), ptr(dk), n, log_n, NULL)
    sync()
    got = host(dd)
    assert np.array_equal(got, po.permute_hc_to_canonic_brt(src))
    assert np.array_equal(host(dk), src)


def test_permute_unaligned_buffers():
    # buffers off 16-byte alignment take the one-word-per-lane path
    log_n, n = 12, 1 << 12
    src = rnd(n, 95)
    ds, dd = dev(np.concatenate([[7], src]).astype(np.uint32)), empty(n + 1)
    call("vx_permute_hc_to_canonic_brt", C.c_void_p(ds.ptr + 4), C.c_void_p(dd.ptr + 4), n, log_n, NULL)
    sync()
    assert np.array_equal(host(dd)[1:], po.permute_hc_to_canonic_brt(src))


def test_permute_rejects_bad_arguments():
    d1, d2 = empty(1 << 8), empty(1 << 8)
    with pytest.raises(VortexError):  # n not 2^log_n
        call("vx_permute_hc_to_canonic_brt", ptr(d1), ptr(d2), 200, 8, NULL)
    with pytest.raises(VortexError):  # in place
        call("vx_permute_canonic_brt_to_hc_natural", ptr(d1), ptr(d1), 256, 8, NULL)


def test_permute_then_stwo_interpolate():
    # a polynomial's values at the half-coset points, natural order, through
    # the permutation and the stwo inverse NTT give back its coefficients
    from vortexstark_amd import stwo
    from vortexstark_amd.device import DeviceBuffer
    log_n = 12
    coeffs = rnd(1 << log_n, 97)
    cache = stwo.StwoNttCache(log_n)
    d = DeviceBuffer.from_host(coeffs)
    stwo.evaluate_stwo(d, cache)                         # canonic bit-reversed values
    hc = stwo.permute_canonic_brt_to_hc_natural(d, log_n)  # half-coset natural order
    back = stwo.permute_hc_to_canonic_brt(hc, log_n)
    stwo.interpolate_stwo(back, cache)
    vx._lib.call("vx_device_sync")
    assert np.array_equal(back.to_host(), coeffs)
    # the natural-order values are f at half_coset(log_n).at(k)
    hcs = po.coset(po.HALF_COSET, log_n)
    vals = hc.to_host()
    for k in (0, 1, 2, 3, 1000, 4095):
        x, y = po.coset_at(hcs, k)
        assert po.stwo_eval_at_point(coeffs, (x, 0, 0, 0), (y, 0, 0, 0)) == (int(vals[k]), 0, 0, 0), k


# ---- the OODS step (src/oods.rs) on a device channel

def _chan_words(ch):
    return np.concatenate([np.frombuffer(ch.state, np.uint32), [ch.c.n_draws]]).astype(np.uint32)


@pytest.mark.parametrize("trace_log,logs", [(0, [3]), (5, [0, 1, 5, 5]), (20, [14, 16, 18, 20, 20]),
                                            (31, [12, 22]), (8, [])])
def test_oods_sample(trace_log, logs):
    print(f"oods_sample trace_log={trace_log} logs={logs}", flush=True)
    cols = [rnd(1 << L, 200 + i) for i, L in enumerate(logs)]
    ch = po.PyChannel()
    ch.mix_digest([trace_log, 8, 6, 7, 5, 3, 0, 9])
    d_ch = dev(_chan_words(ch))
    dcols = [dev(c) for c in cols]
    d_pts, d_vals = empty(16), empty(max(1, 8 * len(cols)))
    ptrs = (C.c_void_p * max(1, len(cols)))(*[c.ptr for c in dcols])
    lg = (C.c_uint32 * max(1, len(cols)))(*logs)
    call("vx_oods_sample", ptr(d_ch), ptrs, lg, len(cols), trace_log, ptr(d_pts), ptr(d_vals), NULL)
    pts, vals = ch.oods_sample(cols, trace_log)
    assert np.array_equal(host(d_pts), pts)
    if cols:
        assert np.array_equal(host(d_vals)[:8 * len(cols)], vals)
    assert np.array_equal(host(d_ch), _chan_words(ch))


def test_oods_sample_mirror():
    from vortexstark_amd import stwo
    from vortexstark_amd.device import DeviceBuffer
    cols = [rnd(1 << L, 300 + L) for L in (10, 13)]
    ch = po.PyChannel()
    ch.mix_digest([1, 2, 3, 4, 5, 6, 7, 8])
    d_ch = DeviceBuffer.from_host(_chan_words(ch))
    pts, vals = stwo.oods_sample(d_ch, [DeviceBuffer.from_host(c) for c in cols], [10, 13], 13)
    want_pts, want_vals = ch.oods_sample(cols, 13)
    assert np.array_equal(pts, want_pts) and np.array_equal(vals, want_vals)
    assert np.array_equal(d_ch.to_host(), _chan_words(ch))
