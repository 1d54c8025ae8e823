"""GKR operations of the stwo plugin on the GPU (vx_gkr_*, replacing
cuda_gkr_*, src/cuda/ffi.rs:1129-1179) against the oracle, bit-exact:
every layer kind, even sizes that are not powers of two, columns off
16-byte alignment (the word-access kernels), folding in place, the eq table
pass by pass and in one launch, and the round sums reduced to one block."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from gpu_util import dev, empty, host, ptr, sync
from vortexstark_amd._lib import VortexError, call

pytestmark = pytest.mark.gpu
NULL = None
P = po.P


def rnd(n, seed):
    return np.random.default_rng(seed).integers(0, P, n, dtype=np.uint32)


def at(buf, off_words):
    """Device pointer `off_words` words into buf."""
    return C.c_void_p(buf.ptr + 4 * off_words)


SIZES = [2, 4, 1000, 1 << 16, 1 << 22]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("secure", [False, True])
def test_fix_first_variable(n, secure):
    print("gkr fix_first", n, secure, flush=True)
    col = rnd(4 * n if secure else n, n + secure)
    r = rnd(4, 5 + n)
    want = po.gkr_fix_first_variable(col, r, secure)
    d, dr, out = dev(col), dev(r), empty(2 * n)
    name = "vx_gkr_fix_first_variable_secure" if secure else "vx_gkr_fix_first_variable_base"
    call(name, ptr(d), ptr(out), ptr(dr), n, NULL)
    assert np.array_equal(host(out), want)


@pytest.mark.parametrize("secure", [False, True])
def test_fix_first_variable_unaligned_and_in_place(secure):
    n = 4096
    col = rnd(4 * n if secure else n, 77)
    r = rnd(4, 78)
    want = po.gkr_fix_first_variable(col, r, secure)
    # one word off 16-byte alignment on both sides
    d = dev(np.concatenate([np.zeros(1, np.uint32), col]))
    out, dr = empty(2 * n + 1), dev(r)
    name = "vx_gkr_fix_first_variable_secure" if secure else "vx_gkr_fix_first_variable_base"
    call(name, at(d, 1), at(out, 1), ptr(dr), n, NULL)
    assert np.array_equal(host(out)[1:], want)
    if secure:  # the plugin's comment allows writing the first half of the input
        d2 = dev(col)
        call(name, ptr(d2), ptr(d2), ptr(dr), n, NULL)
        assert np.array_equal(host(d2)[:2 * n], want)


@pytest.mark.parametrize("k", [0, 1, 5, 10, 11, 16, 20])
def test_gen_eq_evals_passes_and_one_launch(k):
    print("gkr eq", k, flush=True)
    y = rnd(4 * k, 300 + k) if k else np.zeros(4, np.uint32)
    v = rnd(4, 301 + k)
    want = po.gkr_gen_eq_evals(y[:4 * k], v)
    # the plugin's sequence: init, then passes with y[k-1] .. y[0]
    buf, dv = empty(4 << k), dev(v)
    ys = [dev(y[4 * j:4 * j + 4]) for j in range(k)]
    call("vx_gkr_gen_eq_evals_init", ptr(buf), ptr(dv), NULL)
    for i in range(k):
        call("vx_gkr_gen_eq_evals_pass", ptr(buf), ptr(ys[k - 1 - i]), 1 << i, NULL)
    assert np.array_equal(host(buf), want)
    one, dy = empty(4 << k), dev(y)
    call("vx_gkr_gen_eq_evals", ptr(one), ptr(dy), k, ptr(dv), NULL)
    assert np.array_equal(host(one), want)


def test_gen_eq_evals_one_launch_unaligned():
    k = 12
    y, v = rnd(4 * k, 9), rnd(4, 10)
    want = po.gkr_gen_eq_evals(y, v)
    buf, dy, dv = empty((4 << k) + 1), dev(y), dev(v)
    call("vx_gkr_gen_eq_evals", at(buf, 1), ptr(dy), k, ptr(dv), NULL)
    assert np.array_equal(host(buf)[1:], want)


def layer_arrays(kind, n, seed):
    num = None
    if kind in (po.GKR_GRAND_PRODUCT, po.GKR_LOGUP_GENERIC):
        num = rnd(4 * n, seed)
    elif kind == po.GKR_LOGUP_MULT:
        num = rnd(n, seed)
    den = None if kind == po.GKR_GRAND_PRODUCT else rnd(4 * n, seed + 1)
    return num, den


LAYER_FN = {po.GKR_GRAND_PRODUCT: "vx_gkr_next_layer_grand_product",
            po.GKR_LOGUP_GENERIC: "vx_gkr_next_layer_logup_generic",
            po.GKR_LOGUP_MULT: "vx_gkr_next_layer_logup_mult",
            po.GKR_LOGUP_SINGLES: "vx_gkr_next_layer_logup_singles"}


def run_layer(kind, dn, dd, on, od, n):
    if kind == po.GKR_GRAND_PRODUCT:
        call(LAYER_FN[kind], dn, on, n, NULL)
    elif kind == po.GKR_LOGUP_SINGLES:
        call(LAYER_FN[kind], dd, on, od, n, NULL)
    else:
        call(LAYER_FN[kind], dn, dd, on, od, n, NULL)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_next_layer(kind, n):
    print("gkr layer", kind, n, flush=True)
    num, den = layer_arrays(kind, n, 40 + n + kind)
    wn, wd = po.gkr_next_layer(kind, num, den, n)
    dn = dev(num) if num is not None else None
    dd = dev(den) if den is not None else None
    on, od = empty(2 * n), empty(2 * n)
    run_layer(kind, ptr(dn) if dn else None, ptr(dd) if dd else None, ptr(on), ptr(od), n)
    assert np.array_equal(host(on), wn)
    if wd is not None:
        assert np.array_equal(host(od), wd)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_next_layer_unaligned(kind):
    n = 2048
    num, den = layer_arrays(kind, n, 55 + kind)
    wn, wd = po.gkr_next_layer(kind, num, den, n)
    pad = np.zeros(1, np.uint32)
    dn = dev(np.concatenate([pad, num])) if num is not None else None
    dd = dev(np.concatenate([pad, den])) if den is not None else None
    on, od = empty(2 * n + 1), empty(2 * n + 1)
    run_layer(kind, at(dn, 1) if dn else None, at(dd, 1) if dd else None, at(on, 1), at(od, 1), n)
    assert np.array_equal(host(on)[1:], wn)
    if wd is not None:
        assert np.array_equal(host(od)[1:], wd)


SUM_FN = {po.GKR_GRAND_PRODUCT: "vx_gkr_sum_poly_grand_product",
          po.GKR_LOGUP_GENERIC: "vx_gkr_sum_poly_logup_generic",
          po.GKR_LOGUP_MULT: "vx_gkr_sum_poly_logup_mult",
          po.GKR_LOGUP_SINGLES: "vx_gkr_sum_poly_logup_singles"}


def run_sum(kind, deq, dn, dd, dl, part, n_terms):
    nb = C.c_uint32(0)
    if kind == po.GKR_GRAND_PRODUCT:
        call(SUM_FN[kind], deq, dn, part, n_terms, C.byref(nb), NULL)
    elif kind == po.GKR_LOGUP_SINGLES:
        call(SUM_FN[kind], deq, dd, dl, part, n_terms, C.byref(nb), NULL)
    else:
        call(SUM_FN[kind], deq, dn, dd, dl, part, n_terms, C.byref(nb), NULL)
    return nb.value


@pytest.mark.parametrize("n_terms", [1, 255, 256, 257, 5000, 1 << 20])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_sum_poly(kind, n_terms):
    print("gkr sum", kind, n_terms, flush=True)
    num, den = layer_arrays(kind, 4 * n_terms, 90 + kind + n_terms)
    eq, lam = rnd(4 * n_terms, 91 + n_terms), rnd(4, 92)
    w0, w2 = po.gkr_sum_poly(kind, eq, num, den, lam if kind else None, n_terms)
    deq, dl = dev(eq), dev(lam)
    dn = dev(num) if num is not None else None
    dd = dev(den) if den is not None else None
    n_blocks = (n_terms + 255) // 256  # what the plugin allocates
    part = dev(np.full(8 * n_blocks, 0xFFFFFFFF, np.uint32))
    nb = run_sum(kind, ptr(deq), ptr(dn) if dn else None, ptr(dd) if dd else None, ptr(dl), ptr(part), n_terms)
    assert nb == 1
    got = host(part)[:8]
    assert np.array_equal(got[:4], w0) and np.array_equal(got[4:], w2)


def test_sum_poly_unaligned():
    kind, n_terms = po.GKR_LOGUP_GENERIC, 3000
    num, den = layer_arrays(kind, 4 * n_terms, 5)
    eq, lam = rnd(4 * n_terms, 6), rnd(4, 7)
    w0, w2 = po.gkr_sum_poly(kind, eq, num, den, lam, n_terms)
    pad = np.zeros(1, np.uint32)
    deq, dn, dd, dl = dev(np.concatenate([pad, eq])), dev(np.concatenate([pad, num])), \
        dev(np.concatenate([pad, den])), dev(lam)
    part = empty(8 * ((n_terms + 255) // 256))
    assert run_sum(kind, at(deq, 1), at(dn, 1), at(dd, 1), ptr(dl), ptr(part), n_terms) == 1
    got = host(part)[:8]
    assert np.array_equal(got[:4], w0) and np.array_equal(got[4:], w2)


def test_gkr_rejects_bad_arguments():
    d, out, r = empty(16), empty(64), empty(4)
    with pytest.raises(VortexError, match="even"):
        call("vx_gkr_fix_first_variable_secure", ptr(d), ptr(out), ptr(r), 3, NULL)
    with pytest.raises(VortexError, match="null"):
        call("vx_gkr_next_layer_logup_generic", NULL, ptr(d), ptr(out), ptr(out), 4, NULL)
    nb = C.c_uint32(0)
    with pytest.raises(VortexError, match="n_terms"):
        call("vx_gkr_sum_poly_grand_product", ptr(d), ptr(d), ptr(out), 0, C.byref(nb), NULL)
    with pytest.raises(VortexError, match="31 variables"):
        call("vx_gkr_gen_eq_evals", ptr(out), ptr(d), 32, ptr(r), NULL)
    sync()
