"""stwo-basis circle NTT (the stwo plugin's PolyOps) on the GPU, through the
C-ABI, against the oracle (pinned by tests/test_oracle_stwo.py).

Every output is bit-identical to the oracle's: the device-built twiddle
buffers, evaluate / interpolate from n = 4 (the reference's pad-word case)
through the register-radix sizes, the batched calls over separately allocated
columns (one launch per pass) and the strided asynchronous variant. At 2^24
the checks are size-independent: the roundtrip and the direct sum at sampled
CircleDomain points."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import vortexstark_amd as vx
from gpu_util import Pin, dev, empty, host, ptr, sync
from vortexstark_amd import stwo
from vortexstark_amd._lib import VortexError, call
from vortexstark_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu
P = po.P
NULL = None


def rnd(n, seed):
    return np.random.default_rng(seed).integers(0, P, int(n), dtype=np.uint32)


def bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2)


def cols_arg(ptrs):
    return (C.c_void_p * len(ptrs))(*ptrs)


@pytest.fixture(scope="module", autouse=True)
def _lib_loaded():
    vx.load()
    print("stwo ntt tests: library loaded", flush=True)


def dev_twiddles(log_n):
    half = 1 << (log_n - 1)
    tw, itw = empty(half), empty(half)
    call("vx_stwo_twiddles", log_n, ptr(tw), ptr(itw), NULL)
    sync()
    return tw, itw


@pytest.mark.parametrize("log_n", [2, 3, 4, 7, 13, 14, 16, 22])
def test_twiddles_match_oracle(log_n):
    tw, itw = dev_twiddles(log_n)
    fwd, inv = po.stwo_twiddles(log_n)
    assert np.array_equal(host(tw), fwd)
    assert np.array_equal(host(itw), inv)


@pytest.mark.parametrize("log_n", [2, 3, 4, 5, 6, 9, 12, 13, 14, 15, 16, 17, 19, 20, 23])
def test_evaluate_interpolate_match_oracle(log_n):
    print(f"stwo ntt log_n={log_n}", flush=True)
    n = 1 << log_n
    tw, itw = dev_twiddles(log_n)
    fwd, inv = po.stwo_twiddles(log_n)
    x = rnd(n, 100 + log_n)
    d = dev(x)
    call("vx_stwo_ntt_evaluate", ptr(d), ptr(tw), n, NULL)
    ev = host(d)
    assert np.array_equal(ev, po.stwo_evaluate(x, fwd, omp=True))
    y = rnd(n, 200 + log_n)
    d2 = dev(y)
    call("vx_stwo_ntt_interpolate", ptr(d2), ptr(itw), n, NULL)
    assert np.array_equal(host(d2), po.stwo_interpolate(y, inv, omp=True))
    call("vx_stwo_ntt_interpolate", ptr(d), ptr(itw), n, NULL)
    assert np.array_equal(host(d), x)


def test_host_built_twiddles_give_the_same_values():
    # the plugin uploads host-computed twiddles (poly_ops.rs:102-127)
    log_n = 15
    fwd, inv = po.stwo_twiddles(log_n)
    x = rnd(1 << log_n, 7)
    d, tw = dev(x), dev(fwd)
    call("vx_stwo_ntt_evaluate", ptr(d), ptr(tw), 1 << log_n, NULL)
    assert np.array_equal(host(d), po.stwo_evaluate(x, fwd, omp=True))


@pytest.mark.parametrize("log_n,n_cols", [(3, 4), (10, 5), (14, 3), (16, 7), (18, 2)])
def test_batch_over_separate_columns(log_n, n_cols):
    n = 1 << log_n
    fwd, inv = po.stwo_twiddles(log_n)
    tw, itw = dev_twiddles(log_n)
    xs = [rnd(n, 300 + c) for c in range(n_cols)]
    bufs = [dev(x) for x in xs]
    call("vx_stwo_ntt_batch_evaluate", cols_arg([b.ptr for b in bufs]), ptr(tw), n, n_cols, NULL)
    for x, b in zip(xs, bufs):
        assert np.array_equal(host(b), po.stwo_evaluate(x, fwd, omp=True))
    call("vx_stwo_ntt_batch_interpolate", cols_arg([b.ptr for b in bufs]), ptr(itw), n, n_cols, NULL)
    for x, b in zip(xs, bufs):
        assert np.array_equal(host(b), x)


def test_batch_of_more_columns_than_one_grid_row():
    # 70000 four-point columns 16 bytes apart in one allocation: two chunks
    n_cols, n = 70000, 4
    x = rnd(n * n_cols, 9)
    big = dev(x)
    tw, _ = dev_twiddles(2)
    call("vx_stwo_ntt_batch_evaluate", cols_arg([big.ptr + 16 * c for c in range(n_cols)]), ptr(tw), n, n_cols,
         NULL)
    got = host(big).reshape(n_cols, n)
    fwd, _ = po.stwo_twiddles(2)
    for c in list(range(0, n_cols, 997)) + [65534, 65535, 65536, n_cols - 1]:
        assert np.array_equal(got[c], po.stwo_evaluate(x[n * c:n * c + n], fwd)), c


def test_batch_without_columns_is_a_no_op():
    tw, _ = dev_twiddles(4)
    call("vx_stwo_ntt_batch_evaluate", None, ptr(tw), 16, 0, NULL)


def test_strided_async_batch():
    log_n, n_cols = 15, 3
    n, stride = 1 << log_n, (1 << log_n) + 64
    fwd, inv = po.stwo_twiddles(log_n)
    tw, itw = dev_twiddles(log_n)
    x = rnd(stride * n_cols, 12)
    d = dev(x)
    call("vx_stwo_ntt_evaluate_async", ptr(d), ptr(tw), n, n_cols, stride, NULL)
    sync()
    got = host(d)
    for c in range(n_cols):
        col = x[c * stride:c * stride + n]
        assert np.array_equal(got[c * stride:c * stride + n], po.stwo_evaluate(col, fwd, omp=True))
        assert np.array_equal(got[c * stride + n:(c + 1) * stride], x[c * stride + n:(c + 1) * stride])
    call("vx_stwo_ntt_interpolate_async", ptr(d), ptr(itw), n, n_cols, stride, NULL)
    sync()
    assert np.array_equal(host(d), x)


def test_full_size_roundtrip_and_sampled_points():
    log_n = 24
    n = 1 << log_n
    print("stwo ntt log_n=24 roundtrip", flush=True)
    cache = stwo.StwoNttCache(log_n)
    x = rnd(n, 24)
    d = DeviceBuffer.from_host(x)
    stwo.evaluate_stwo(d, cache)
    ev = d.to_host()
    for k in (0, 1, 12345, n - 1):
        zx, zy = po.stwo_domain_point(log_n, bitrev(k, log_n))
        assert po.stwo_eval_at_point(x, (zx, 0, 0, 0), (zy, 0, 0, 0)) == (int(ev[k]), 0, 0, 0), k
    stwo.interpolate_stwo(d, cache)
    assert np.array_equal(d.to_host(), x)


def test_mirror_reference_vector_log8():
    # src/ntt.rs:517-548 through the Python mirror of StwoNttCache / evaluate_stwo
    log_n = 8
    cache = stwo.StwoNttCache(log_n)
    c = np.zeros(1 << log_n, np.uint32)
    c[:4] = [42, 17, 5, 99]
    d = DeviceBuffer.from_host(c)
    stwo.evaluate_stwo(d, cache)
    ev = d.to_host()
    for k in range(1 << log_n):
        zx, zy = po.stwo_domain_point(log_n, bitrev(k, log_n))
        assert po.stwo_eval_at_point(c, (zx, 0, 0, 0), (zy, 0, 0, 0)) == (int(ev[k]), 0, 0, 0)
    cols = [DeviceBuffer.from_host(rnd(1 << log_n, s)) for s in range(3)]
    orig = [b.to_host() for b in cols]
    stwo.evaluate_stwo_batch(cols, cache)
    stwo.interpolate_stwo_batch(cols, cache)
    assert all(np.array_equal(b.to_host(), o) for b, o in zip(cols, orig))


def test_bad_arguments_are_rejected():
    tw, _ = dev_twiddles(4)
    d = dev(rnd(16, 1))
    for n in (2, 1, 24, 0):
        with pytest.raises(VortexError):
            call("vx_stwo_ntt_evaluate", ptr(d), ptr(tw), n, NULL)
    with pytest.raises(VortexError):
        call("vx_stwo_ntt_evaluate", ptr(d), NULL, 16, NULL)
    with pytest.raises(VortexError):  # misaligned column
        call("vx_stwo_ntt_batch_evaluate", cols_arg([d.ptr + 4]), ptr(tw), 16, 1, NULL)
    with pytest.raises(VortexError):
        call("vx_stwo_twiddles", 1, ptr(tw), NULL, NULL)
    # nothing was written by the rejected calls
    assert np.array_equal(host(d), rnd(16, 1))


# ---- evaluation at a point (PolyOps::eval_at_point) and barycentric sums

def eval_gpu(coeffs_dev, factors, n):
    words = 4 if n < (1 << 14) else max(80, n // 1024)  # the documented minimum scratch
    s1, s2 = empty(words), empty(words)
    res = np.zeros(4, np.uint32)
    d_f = dev(factors)
    call("vx_eval_at_point", ptr(coeffs_dev), ptr(d_f), C.c_void_p(res.ctypes.data), n, ptr(s1), ptr(s2), NULL)
    return tuple(int(v) for v in res)


@pytest.mark.parametrize("log_n", [0, 1, 2, 3, 4, 5, 8, 11, 12, 13, 14, 15, 16, 17, 20, 21, 24])
def test_eval_at_point_matches_oracle(log_n):
    print(f"eval_at_point log_n={log_n}", flush=True)
    rng = np.random.default_rng(500 + log_n)
    n = 1 << log_n
    c = rng.integers(0, P, n, dtype=np.uint32)
    zx, zy = (tuple(int(v) for v in rng.integers(0, P, 4)) for _ in range(2))
    f = po.folding_factors(zx, zy, log_n)
    assert eval_gpu(dev(c), f, n) == po.eval_at_point_factors(c, f)


def test_eval_at_point_agrees_with_evaluate():
    # value k of the stwo NTT is f at CircleDomain point bitrev(k)
    log_n = 20
    n = 1 << log_n
    c = rnd(n, 77)
    d = dev(c)
    cache = stwo.StwoNttCache(log_n)
    e = DeviceBuffer.from_host(c)
    stwo.evaluate_stwo(e, cache)
    ev = e.to_host()
    for k in (0, 3, 99999, n - 1):
        x, y = po.stwo_domain_point(log_n, bitrev(k, log_n))
        assert eval_gpu(d, po.folding_factors((x, 0, 0, 0), (y, 0, 0, 0), log_n), n) == (int(ev[k]), 0, 0, 0)


@pytest.mark.parametrize("n,n_blocks", [(0, 1), (1, 1), (255, 1), (256, 1), (1000, 4), (4097, 3),
                                        (1 << 20, 256), (1 << 20, 7), (12345, 300)])
def test_barycentric_partials_match_oracle(n, n_blocks):
    rng = np.random.default_rng(n + n_blocks)
    e = rng.integers(0, P, max(n, 1), dtype=np.uint32)[:n]
    w = rng.integers(0, P, max(4 * n, 4), dtype=np.uint32)[:4 * n]
    de, dw = dev(e if n else np.zeros(1, np.uint32)), dev(w if n else np.zeros(4, np.uint32))
    out = empty(4 * n_blocks)
    call("vx_barycentric_eval", ptr(de), ptr(dw), n, ptr(out), n_blocks, NULL)
    assert np.array_equal(host(out), po.barycentric_partials(e, w, n_blocks))


def test_mirror_eval_at_point_and_barycentric():
    log_n = 14
    n = 1 << log_n
    c = rnd(n, 31)
    zx, zy = (5, 6, 7, 8), (9, 10, 11, 12)
    got = stwo.eval_at_point(DeviceBuffer.from_host(c), n, zx, zy)
    assert got == po.stwo_eval_at_point(c, zx, zy)
    assert stwo.folding_factors(zx, zy, log_n) == po.folding_factors(zx, zy, log_n).tolist()
    e, w = rnd(3000, 32), rnd(12000, 33)
    total = tuple(int(v) for v in po.barycentric_partials(e, w, 1))
    assert stwo.barycentric_eval(DeviceBuffer.from_host(e), DeviceBuffer.from_host(w), 3000) == total
    pin, res = Pin(), np.zeros(4, np.uint32)
    with pytest.raises(VortexError):
        call("vx_eval_at_point", pin(dev(c)), pin(dev(np.zeros(56, np.uint32))), C.c_void_p(res.ctypes.data),
             n - 1, pin(empty(64)), pin(empty(64)), NULL)


@pytest.mark.parametrize("logs,max_log_n", [([0, 1, 2, 3, 5, 10, 13, 14, 15, 17, 22, 14, 2], 22),
                                             ([16, 16, 16, 16], 24), ([4], 4), ([20] * 3 + [23], 23)])
def test_eval_at_point_batch(logs, max_log_n):
    print(f"eval_at_point_batch logs={logs} max={max_log_n}", flush=True)
    rng = np.random.default_rng(sum(logs) + max_log_n)
    zx, zy = (tuple(int(v) for v in rng.integers(0, P, 4)) for _ in range(2))
    cols = [rng.integers(0, P, 1 << L, dtype=np.uint32) for L in logs]
    bufs = [dev(c) for c in cols]
    f = dev(po.folding_factors(zx, zy, max_log_n))
    res = empty(4 * len(logs))
    call("vx_eval_at_point_batch", (C.c_void_p * len(logs))(*[b.ptr for b in bufs]),
         (C.c_uint32 * len(logs))(*logs), len(logs), ptr(f), max_log_n, ptr(res), NULL)
    got = host(res).reshape(-1, 4)
    for k, c in enumerate(cols):
        assert tuple(int(v) for v in got[k]) == po.stwo_eval_at_point(c, zx, zy), (k, logs[k])
    with pytest.raises(VortexError):  # a column larger than the factors cover
        call("vx_eval_at_point_batch", (C.c_void_p * 1)(bufs[0].ptr), (C.c_uint32 * 1)(max_log_n + 1), 1, ptr(f),
             max_log_n, ptr(res), NULL)


@pytest.mark.parametrize("coeff_log,log_n", [(1, 2), (3, 5), (10, 12), (14, 16), (18, 20), (20, 20), (22, 24)])
def test_stwo_evaluate_lde(coeff_log, log_n):
    print(f"stwo lde {coeff_log} -> {log_n}", flush=True)
    n = 1 << log_n
    c = rnd(1 << coeff_log, 400 + coeff_log)
    tw, _ = dev_twiddles(log_n)
    d_c, d_e = dev(c), empty(n)
    call("vx_stwo_ntt_evaluate_lde", ptr(d_e), ptr(d_c), 1 << coeff_log, ptr(tw), n, NULL)
    sync()
    padded = np.zeros(n, np.uint32)
    padded[:c.size] = c
    fwd, _ = po.stwo_twiddles(log_n)
    assert np.array_equal(host(d_e), po.stwo_evaluate(padded, fwd, omp=True))
    assert np.array_equal(host(d_c), c)  # the coefficients are only read


@pytest.mark.parametrize("coeff_log,log_n,n_cols", [(12, 14, 3), (16, 18, 5), (19, 21, 2)])
def test_stwo_evaluate_lde_batch(coeff_log, log_n, n_cols):
    n = 1 << log_n
    cs = [rnd(1 << coeff_log, 500 + k) for k in range(n_cols)]
    dcs = [dev(c) for c in cs]
    des = [empty(n) for _ in range(n_cols)]
    tw, _ = dev_twiddles(log_n)
    call("vx_stwo_ntt_evaluate_lde_batch", cols_arg([d.ptr for d in des]), cols_arg([d.ptr for d in dcs]),
         1 << coeff_log, n_cols, ptr(tw), n, NULL)
    fwd, _ = po.stwo_twiddles(log_n)
    for c, d in zip(cs, des):
        padded = np.zeros(n, np.uint32)
        padded[:c.size] = c
        assert np.array_equal(host(d), po.stwo_evaluate(padded, fwd, omp=True))
