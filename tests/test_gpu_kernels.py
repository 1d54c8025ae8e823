"""Per-kernel parity of the HIP path against the CPU oracle, through the C-ABI.

Inputs are seeded; every output must be bit-identical to the oracle's.
Sizes run from the smallest shapes the reference allows up to 2^22-2^24 (the
oracle finishes these in seconds)."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import vortexstark_amd as vx
from gpu_util import d2h, dev, empty, host, ptr, sync
from vortexstark_amd._lib import call

pytestmark = pytest.mark.gpu
P = po.P
NULL = None


def rnd(n, seed):
    return np.random.default_rng(seed).integers(0, P, int(n), dtype=np.uint32)


def hp(a):
    """host pointer of a numpy array"""
    return C.c_void_p(a.ctypes.data)


@pytest.fixture(scope="module", autouse=True)
def _lib_loaded():
    vx.load()
    print("gpu kernel parity tests: library loaded", flush=True)


@pytest.mark.parametrize("log_n,a,b", [(4, 1, 1), (5, 3, 7), (10, 123, 456), (16, 1, 2 ** 30), (22, P - 1, P - 2)])
def test_fibonacci_trace(log_n, a, b):
    out = empty(1 << log_n)
    call("vx_fibonacci_trace", a, b, ptr(out), log_n, NULL)
    assert np.array_equal(host(out), po.fibonacci_trace(a, b, log_n))


@pytest.mark.parametrize("n", [1, 48, 1000, 4096 + 16])
def test_coset_points_any_length(n):
    # lengths that are not powers of two: multiples of 16 take the 16-per-lane
    # kernel, the others the run-per-lane one
    c = po.coset(po.HALF_ODDS, 14)
    xs, ys = empty(n), empty(n)
    call("vx_compute_coset_points", int(c[0]), int(c[1]), int(c[2]), int(c[3]), ptr(xs), ptr(ys), n, NULL)
    hx, hy = host(xs), host(ys)
    for i in sorted({0, min(1, n - 1), n // 3, n // 2, n - 1}):
        assert (int(hx[i]), int(hy[i])) == po.coset_at(c, i), i


@pytest.mark.parametrize("kind,log_n", [(po.HALF_COSET, 3), (po.HALF_COSET, 11), (po.HALF_ODDS, 10),
                                        (po.SUBGROUP, 12), (po.HALF_COSET, 20)])
def test_coset_points_and_squash(kind, log_n):
    c = po.coset(kind, log_n)
    n = 1 << log_n
    xs, ys = empty(n), empty(n)
    call("vx_compute_coset_points", int(c[0]), int(c[1]), int(c[2]), int(c[3]), ptr(xs), ptr(ys), n, NULL)
    hx, hy = host(xs), host(ys)
    line, circ, offs, sizes = po.twiddles(c)
    assert np.array_equal(hy[:n // 2], circ[:n // 2])
    # one extract_and_squash layer reproduces line layer 0 and feeds layer 1
    tw, sq = empty(n // 2), empty(n // 2)
    call("vx_extract_and_squash", ptr(xs), ptr(tw), ptr(sq), n // 2, NULL)
    assert np.array_equal(host(tw), line[:n // 2])
    tw2, sq2 = empty(n // 4), empty(n // 4)
    call("vx_extract_and_squash", ptr(sq), ptr(tw2), ptr(sq2), n // 4, NULL)
    assert np.array_equal(host(tw2), line[offs[1]:offs[1] + sizes[1]])
    sq3 = empty(n // 2)
    call("vx_squash_x", ptr(xs), ptr(sq3), n // 2, NULL)
    assert np.array_equal(host(sq3), host(sq))
    assert np.array_equal(hx[:4], [po.coset_at(c, i)[0] for i in range(4)])


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 1 << 20])
def test_batch_inverse(n):
    x = rnd(n, n)
    x[::7] = 0
    d = dev(x)
    out = empty(n)
    call("vx_batch_inverse_m31", ptr(d), ptr(out), n, NULL)
    assert np.array_equal(host(out), po.batch_inverse(x))
    call("vx_batch_inverse_m31", ptr(d), ptr(d), n, NULL)  # in place
    assert np.array_equal(host(d), po.batch_inverse(x))


@pytest.mark.parametrize("kind,log_n,use_y", [(po.HALF_COSET, 4, True), (po.HALF_COSET, 5, False),
                                              (po.HALF_ODDS, 6, True), (po.HALF_COSET, 18, True),
                                              (po.HALF_ODDS, 9, False), (po.HALF_ODDS, 20, False),
                                              (po.SUBGROUP, 12, True)])
def test_fold_twiddle_sources(kind, log_n, use_y):
    c = po.coset(kind, log_n)
    n = 1 << log_n
    out = empty(n // 2)
    call("vx_compute_fold_twiddle_sources", int(c[0]), int(c[1]), int(c[2]), int(c[3]), ptr(out), n, log_n,
         1 if use_y else 0, NULL)
    src = host(out)
    inv = empty(n // 2)
    call("vx_batch_inverse_m31", ptr(out), ptr(inv), n // 2, NULL)
    assert np.array_equal(host(inv), po.fold_twiddles(c, use_y))
    assert src[0] == (po.coset_at(c, 0)[1] if use_y else po.coset_at(c, 0)[0])


def _tables(c):
    line, circ, offs, sizes = po.twiddles(c)
    return dev(line), dev(circ), np.ascontiguousarray(offs, np.uint32), np.ascontiguousarray(sizes, np.uint32)


def _itables(c):
    line, circ, offs, sizes = po.inverse_twiddles(c)
    return dev(line), dev(circ), np.ascontiguousarray(offs, np.uint32), np.ascontiguousarray(sizes, np.uint32)


NTT_LOGS = [2, 3, 4, 5, 8, 12, 13, 14, 15, 16, 18, 21, 24]


@pytest.mark.parametrize("log_n", NTT_LOGS)
def test_ntt_evaluate(log_n):
    c = po.coset(po.HALF_COSET, log_n)
    v = rnd(1 << log_n, 100 + log_n)
    line, circ, offs, sizes = _tables(c)
    d = dev(v)
    call("vx_circle_ntt_evaluate", ptr(d), ptr(line), ptr(circ), hp(offs), hp(sizes), log_n - 1, 1 << log_n, NULL)
    assert np.array_equal(host(d), po.evaluate(v, c))


@pytest.mark.parametrize("log_n", NTT_LOGS)
def test_ntt_interpolate(log_n):
    c = po.coset(po.HALF_COSET, log_n)
    v = rnd(1 << log_n, 200 + log_n)
    line, circ, offs, sizes = _itables(c)
    d = dev(v)
    call("vx_circle_ntt_interpolate", ptr(d), ptr(line), ptr(circ), hp(offs), hp(sizes), log_n - 1, 1 << log_n,
         NULL)
    assert np.array_equal(host(d), po.interpolate(v, c))


@pytest.mark.parametrize("log_n,n_cols", [(2, 3), (5, 1), (10, 7), (16, 4), (20, 2)])
@pytest.mark.parametrize("variant", ["device_array", "host_array"])
def test_ntt_batch_evaluate_and_interpolate(log_n, n_cols, variant):
    # cuda_circle_ntt_evaluate_batch / _interpolate_batch (ffi.rs:281-301,
    # device pointer array) and the FORGE entry points (ffi.rs:407-428, host
    # array): every column equals the oracle's single-column transform
    c = po.coset(po.HALF_COSET, log_n)
    n = 1 << log_n
    vals = [rnd(n, 900 + 31 * log_n + k) for k in range(n_cols)]
    cols = [dev(v) for v in vals]
    suffix = "" if variant == "device_array" else "_host"
    arr = _dev_ptr_array(cols) if variant == "device_array" else None
    colptr = ptr(arr) if arr is not None else _col_ptr_array(cols)
    fl, fc, offs, sizes = _tables(c)
    call("vx_circle_ntt_evaluate_batch" + suffix, colptr, ptr(fl), ptr(fc), hp(offs), hp(sizes), log_n - 1, n,
         n_cols, NULL)
    for d, v in zip(cols, vals):
        assert np.array_equal(host(d), po.evaluate(v, c))
    il, ic, ioffs, isizes = _itables(c)
    evals = [host(d) for d in cols]
    call("vx_circle_ntt_interpolate_batch" + suffix, colptr, ptr(il), ptr(ic), hp(ioffs), hp(isizes), log_n - 1,
         n, n_cols, NULL)
    for d, e, v in zip(cols, evals, vals):
        got = host(d)
        assert np.array_equal(got, po.interpolate(e, c))
        assert np.array_equal(got, v)


def test_ntt_roundtrip_2p26():
    # size-independent property at the LDE size of the log_n = 24 proof
    log_n = 26
    c = po.coset(po.HALF_COSET, log_n)
    fl, fc, offs, sizes = _tables(c)
    il, ic, _, _ = _itables(c)
    v = rnd(1 << log_n, 26)
    d = dev(v)
    call("vx_circle_ntt_evaluate", ptr(d), ptr(fl), ptr(fc), hp(offs), hp(sizes), log_n - 1, 1 << log_n, NULL)
    assert not np.array_equal(host(d), v)
    call("vx_circle_ntt_interpolate", ptr(d), ptr(il), ptr(ic), hp(offs), hp(sizes), log_n - 1, 1 << log_n, NULL)
    assert np.array_equal(host(d), v)


@pytest.mark.parametrize("log_n,blowup", [(3, 2), (4, 2), (8, 1), (12, 2), (14, 2), (16, 3), (20, 2)])
def test_lde_matches_zero_pad_then_evaluate(log_n, blowup):
    le = log_n + blowup
    c = po.eval_domain(le)
    coef = rnd(1 << log_n, le)
    padded = np.zeros(1 << le, np.uint32)
    padded[:1 << log_n] = coef
    line, circ, offs, _ = _tables(c)
    dc, out = dev(coef), empty(1 << le)
    call("vx_circle_ntt_evaluate_lde", ptr(out), ptr(dc), 1 << log_n, ptr(line), ptr(circ), hp(offs), le - 1,
         1 << le, NULL)
    assert np.array_equal(host(out), po.evaluate(padded, c))


@pytest.mark.parametrize("log_n,li", [(6, 0), (6, 3), (10, 9), (16, 7)])
@pytest.mark.parametrize("forward", [1, 0])
def test_ntt_single_layer(log_n, li, forward):
    c = po.coset(po.HALF_COSET, log_n)
    line, circ, offs, _ = po.twiddles(c)
    tw = circ if li == 0 else line[offs[li - 1]:]
    v = rnd(1 << log_n, li + 7 * forward)
    d, dt = dev(v), dev(tw)
    call("vx_circle_ntt_layer", ptr(d), ptr(dt), li, 1 << log_n, forward, NULL)
    assert np.array_equal(host(d), po.ntt_layer(v, tw, li, forward))


@pytest.mark.parametrize("log_n", [1, 5, 12, 20])
def test_bit_reverse(log_n):
    v = rnd(1 << log_n, log_n)
    d = dev(v)
    call("vx_bit_reverse_m31", ptr(d), log_n, NULL)
    assert np.array_equal(host(d), po.bit_reverse(v))
    q = rnd(4 << log_n, log_n + 1)
    dq, out = dev(q), empty(4 << log_n)
    call("vx_bit_reverse_qm31", ptr(dq), ptr(out), 1 << log_n, log_n, NULL)
    idx = np.array([int(format(i, f"0{log_n}b")[::-1], 2) for i in range(1 << log_n)])
    want = np.zeros_like(q).reshape(-1, 4)
    want[idx] = q.reshape(-1, 4)
    assert np.array_equal(host(out), want.reshape(-1))


@pytest.mark.parametrize("n", [8, 1000, 1 << 20])
def test_zero_pad_and_quotient(n):
    t = rnd(n, n)
    alpha = rnd(4, n + 1)
    dt = dev(t)
    outs = [empty(n) for _ in range(4)]
    call("vx_fibonacci_quotient", ptr(dt), *[ptr(o) for o in outs], hp(alpha), n, NULL)
    want = po.quotient(t, alpha)
    for o, w in zip(outs, want):
        assert np.array_equal(host(o), w)
    # chunked variant over the tail, including the wrap-around rows
    off, chunk = n // 2, n - n // 2
    outs2 = [empty(chunk) for _ in range(4)]
    call("vx_fibonacci_quotient_chunk", ptr(dt), *[ptr(o) for o in outs2], hp(alpha), off, chunk, n, NULL)
    for o, w in zip(outs2, want):
        assert np.array_equal(host(o), w[off:])
    dst = empty(2 * n)
    call("vx_zero_pad", ptr(dt), ptr(dst), n, 2 * n, NULL)
    assert np.array_equal(host(dst), np.concatenate([t, np.zeros(n, np.uint32)]))


@pytest.mark.parametrize("half_n", [4, 1000, 1 << 19])
def test_folds(half_n):
    cols = [rnd(2 * half_n, half_n + k) for k in range(4)]
    tw = rnd(half_n, 5)
    alpha = rnd(4, 6)
    a2 = po.qm31_mul(alpha, alpha)
    dcols, dtw = [dev(c) for c in cols], dev(tw)
    outs = [empty(half_n) for _ in range(4)]
    call("vx_fold_line_soa", *[ptr(c) for c in dcols], ptr(dtw), *[ptr(o) for o in outs], hp(alpha), half_n, NULL)
    for o, w in zip(outs, po.fold_line(cols, tw, alpha)):
        assert np.array_equal(host(o), w)
    dst = [rnd(half_n, 50 + k) for k in range(4)]
    ddst = [dev(d) for d in dst]
    call("vx_fold_circle_into_line_soa", *[ptr(d) for d in ddst], *[ptr(c) for c in dcols], ptr(dtw), hp(alpha),
         hp(np.ascontiguousarray(a2)), half_n, NULL)
    for o, w in zip(ddst, po.fold_circle_into_line(dst, cols, tw, alpha, a2)):
        assert np.array_equal(host(o), w)


def _col_ptr_array(dcols):
    return (C.c_void_p * len(dcols))(*[t.data_ptr() for t in dcols])


def _dev_ptr_array(dcols):
    """The column pointers uploaded to a device array, as the reference's
    callers do before cuda_merkle_hash_leaves / cuda_merkle_tiled_generic
    (src/merkle.rs:32-34): 64-bit words in device memory."""
    words = np.array([t.data_ptr() for t in dcols], dtype=np.uint64).view(np.uint32)
    return dev(words)


@pytest.mark.parametrize("ncols,n", [(1, 64), (4, 1000), (7, 4096), (16, 256), (19, 128)])
def test_merkle_leaves_and_nodes(ncols, n):
    cols = [rnd(n, ncols * 10 + k) for k in range(ncols)]
    dcols = [dev(c) for c in cols]
    want = po.leaf_hashes(cols[:16])
    darr = _dev_ptr_array(dcols)
    hashes = empty(8 * n)
    call("vx_merkle_hash_leaves", ptr(darr), ptr(hashes), ncols, n, NULL)
    sync()
    assert np.array_equal(host(hashes).reshape(-1, 8), want)
    hashes2 = empty(8 * n)
    call("vx_merkle_hash_leaves_host", _col_ptr_array(dcols), ptr(hashes2), ncols, n, NULL)
    sync()
    assert np.array_equal(host(hashes2).reshape(-1, 8), want)
    par = empty(8 * (n // 2))
    call("vx_merkle_hash_nodes", ptr(hashes), ptr(par), n // 2, NULL)
    assert np.array_equal(host(par).reshape(-1, 8), po.hash_nodes(want))


@pytest.mark.parametrize("n", [1, 64, 1000, 1 << 16])
def test_merkle_hash_leaves_single_and_quad(n):
    # cuda_merkle_hash_leaves_forge_single / _quad (ffi.rs:797-815): the
    # 1- and 4-column leaf rules of cuda_merkle_hash_leaves
    cols = [rnd(n, 700 + n + k) for k in range(4)]
    dcols = [dev(c) for c in cols]
    h1, h4 = empty(8 * n), empty(8 * n)
    call("vx_merkle_hash_leaves_single", ptr(dcols[0]), ptr(h1), n, NULL)
    call("vx_merkle_hash_leaves_quad", *[ptr(c) for c in dcols], ptr(h4), n, NULL)
    sync()
    assert np.array_equal(host(h1).reshape(-1, 8), po.leaf_hashes(cols[:1]))
    assert np.array_equal(host(h4).reshape(-1, 8), po.leaf_hashes(cols))


@pytest.mark.parametrize("n_nodes", [1, 2, 64, 512, 1024])
def test_merkle_reduce_to_root(n_nodes):
    nodes = rnd(8 * n_nodes, n_nodes)
    d, root = dev(nodes), empty(8)
    call("vx_merkle_reduce_to_root", ptr(d), ptr(root), n_nodes, NULL)
    level = nodes.reshape(-1, 8)
    while len(level) > 1:
        level = po.hash_nodes(level)
    assert np.array_equal(host(root), level[0])


@pytest.mark.parametrize("n", [4, 8, 256, 2048])
def test_merkle_small_soa4_and_merge(n):
    cols = [rnd(n, 300 + n + k) for k in range(4)]
    dcols = [dev(c) for c in cols]
    root = empty(8)
    call("vx_merkle_commit_small_soa4", *[ptr(c) for c in dcols], ptr(root), n, NULL)
    assert np.array_equal(host(root), po.merkle_root(cols))
    par = empty(8 * (n // 2))
    call("vx_merkle_hash_leaves_merge_soa4", *[ptr(c) for c in dcols], ptr(par), n, NULL)
    assert np.array_equal(host(par).reshape(-1, 8), po.hash_nodes(po.leaf_hashes(cols)))


def _tile_roots(cols, n):
    """root of every 1024-leaf tile (the oracle's tree, tile by tile)"""
    out = []
    for t in range(n // 1024):
        level = po.leaf_hashes([c[1024 * t:1024 * (t + 1)] for c in cols])
        for _ in range(10):
            level = po.hash_nodes(level)
        out.append(level[0])
    return np.array(out, dtype=np.uint32)


@pytest.mark.parametrize("ncols,n", [(1, 1 << 10), (1, 1 << 16), (4, 1 << 12), (4, 1 << 18), (3, 1 << 11),
                                     (2, 3 * 1024), (5, 7 * 1024), (4, 5 * 1024), (17, 13 * 1024)])
def test_merkle_tiled(ncols, n):
    # cuda_merkle_tiled_generic accepts any multiple of 1024 leaves
    # (blake2s.cu:631-643: n_leaves / 1024 blocks)
    cols = [rnd(n, 400 + n + k) for k in range(ncols)]
    dcols = [dev(c) for c in cols]
    want = _tile_roots(cols[:16], n)
    variants = ["generic", "generic_host"] + (["soa4"] if ncols == 4 else [])
    darr = _dev_ptr_array(dcols)
    for v in variants:
        roots = empty(8 * (n // 1024))
        if v == "soa4":
            call("vx_merkle_tiled_soa4", *[ptr(c) for c in dcols], ptr(roots), n, NULL)
        elif v == "generic":
            call("vx_merkle_tiled_generic", ptr(darr), ptr(roots), ncols, n, NULL)
        else:
            call("vx_merkle_tiled_generic_host", _col_ptr_array(dcols), ptr(roots), ncols, n, NULL)
        sync()
        assert np.array_equal(host(roots).reshape(-1, 8), want), v


def test_merkle_tiled_rejects_partial_tiles():
    d = dev(rnd(2048, 1))
    roots = empty(16)
    with pytest.raises(vx.VortexError, match="multiple of 1024"):
        call("vx_merkle_tiled_generic_host", _col_ptr_array([d]), ptr(roots), 1, 1536, NULL)


def _tree_levels(leaves):
    """level l of the tree over the given leaf hashes, l = 0 .. log n"""
    levels = [leaves]
    while len(levels[-1]) > 1:
        levels.append(po.hash_nodes(levels[-1]))
    return levels


def _check_commit(d_levels, d_root, levels, lo):
    got = host(d_levels)
    off = 0
    for lv in levels[lo:]:
        n = 8 * len(lv)
        assert np.array_equal(got[off:off + n].reshape(-1, 8), lv)
        off += n
    assert off == got.size
    assert np.array_equal(host(d_root), levels[-1][0])


def _levels_buf(log_n, lo):
    return empty(vx.load().vx_merkle_levels_words(log_n, lo))


@pytest.mark.parametrize("ncols,log_n,lo", [(1, 12, 4), (3, 15, 4), (16, 18, 6), (2, 22, 10)])
def test_merkle_commit_tiled(ncols, log_n, lo):
    cols = [rnd(1 << log_n, 700 + log_n + k) for k in range(ncols)]
    dcols = [dev(c) for c in cols]
    lv, root = _levels_buf(log_n, lo), empty(8)
    call("vx_merkle_commit_tiled", _col_ptr_array(dcols), ncols, log_n, lo, ptr(lv), ptr(root), NULL)
    _check_commit(lv, root, _tree_levels(po.leaf_hashes(cols)), lo)


@pytest.mark.parametrize("log_n", [10, 16, 21])
def test_quotient_and_circle_fold_commits(log_n):
    n = 1 << log_n
    ev, tw = rnd(n, 900 + log_n), rnd(n // 2, 901)
    aq, af = rnd(4, 902), rnd(4, 903)
    d_ev, d_tw = dev(ev), dev(tw)
    lv, root = _levels_buf(log_n, 4), empty(8)
    call("vx_quotient_commit", ptr(d_ev), hp(aq), log_n, 4, ptr(lv), ptr(root), NULL)
    q = po.quotient(ev, aq)
    _check_commit(lv, root, _tree_levels(po.leaf_hashes(q)), 4)
    outs = [empty(n // 2) for _ in range(4)]
    lv2, root2 = _levels_buf(log_n - 1, 4), empty(8)
    call("vx_circle_fold_commit", ptr(d_ev), hp(aq), hp(af), ptr(d_tw), *[ptr(o) for o in outs], log_n, 4,
         ptr(lv2), ptr(root2), NULL)
    af2 = np.ascontiguousarray(po.qm31_mul(af, af))
    want = po.fold_circle_into_line([np.zeros(n // 2, np.uint32) for _ in range(4)], q, tw, af, af2)
    for o, w in zip(outs, want):
        assert np.array_equal(host(o), w)
    _check_commit(lv2, root2, _tree_levels(po.leaf_hashes(want)), 4)


@pytest.mark.parametrize("log_half", [9, 14, 19])
def test_fold_line_commit(log_half):
    half = 1 << log_half
    cols = [rnd(2 * half, 950 + k) for k in range(4)]
    tw, alpha = rnd(half, 960), rnd(4, 961)
    dcols, dtw = [dev(c) for c in cols], dev(tw)
    outs = [empty(half) for _ in range(4)]
    lv, root = _levels_buf(log_half, 4), empty(8)
    call("vx_fold_line_commit", *[ptr(c) for c in dcols], ptr(dtw), *[ptr(o) for o in outs], hp(alpha), log_half, 4,
         ptr(lv), ptr(root), NULL)
    want = po.fold_line(cols, tw, alpha)
    for o, w in zip(outs, want):
        assert np.array_equal(host(o), w)
    _check_commit(lv, root, _tree_levels(po.leaf_hashes(want)), 4)


def test_fused_commits_reject_bad_shapes():
    lv, root = empty(64), empty(8)
    col = dev(rnd(1 << 10, 1))
    with pytest.raises(vx.VortexError, match="bad tree shape"):
        call("vx_merkle_commit_tiled", _col_ptr_array([col]), 1, 10, 11, ptr(lv), ptr(root), NULL)
    with pytest.raises(vx.VortexError, match="1..16 columns"):
        call("vx_merkle_commit_tiled", _col_ptr_array([col]), 0, 10, 4, ptr(lv), ptr(root), NULL)
    assert vx.load().vx_merkle_levels_words(10, 11) == 0
    sync()


@pytest.mark.parametrize("bits", [1, 8, 12, 16, 20])
def test_grind_smallest_nonce(bits):
    ch = po.PyChannel()
    ch.mix_digest([bits, 1, 2, 3, 4, 5, 6, 7])
    prefix = np.zeros(8, np.uint32)
    st = np.frombuffer(ch.state, dtype=np.uint32)
    # prefix digest = Blake2s(0x12345678 || 0^12 || state || bits)  (src/channel.rs:140-150)
    msg = np.zeros(13, np.uint32)
    msg[0] = 0x12345678
    msg[4:12] = st
    msg[12] = bits
    prefix[:] = np.frombuffer(po.blake2s(msg.tobytes()), dtype=np.uint32)
    dp, res = dev(prefix), empty(2)
    call("vx_grind_pow_min", ptr(dp), bits, ptr(res), NULL)
    got = host(res)
    nonce = int(got[0]) | (int(got[1]) << 32)
    assert nonce == ch.grind(bits)
    # reference-style batch over [0, nonce + 1) gives the same minimum
    res2 = dev(np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32))
    call("vx_grind_pow", ptr(dp), ptr(res2), bits, 0, nonce + 1, NULL)
    got2 = host(res2)
    assert int(got2[0]) | (int(got2[1]) << 32) == nonce
