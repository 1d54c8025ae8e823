"""The Cairo VM AIR quotient on the GPU (vx_cairo_quotient / _slab / _chunk,
vx_compute_vanishing[_inv], vx_axpy_m31 through the C-ABI) against the
oracle (or_cairo_quotient, pinned by tests/test_oracle_cairo.py): random
inputs at several sizes, valid Cairo tables (zero quotient), slabs, the
chunked form, the fused quotient + Merkle commitment, the vanishing
polynomial of the trace domain, and the committed oracle digest of a
Cairo-shaped commitment (lo / hi / dict trace trees + quotient tree)."""
import random

import numpy as np
import pytest

import pyoracle as po
import pyref
import pyref_cairo as pc
import vortexstark_amd as vx
from gpu_util import Pin, empty, host, sync

pytestmark = pytest.mark.gpu
P = po.P
call = vx._lib.call


def rand_inputs(seed, n):
    rng = np.random.default_rng(seed)
    cols = [rng.integers(0, P, n, dtype=np.uint32) for _ in range(34)]
    for c in list(range(5, 20)) + [33]:
        cols[c] = rng.integers(0, 2, n, dtype=np.uint32)
    inter = [rng.integers(0, P, n, dtype=np.uint32) for _ in range(32)]
    alpha = rng.integers(0, P, 160, dtype=np.uint32)
    chal = rng.integers(0, P, 24, dtype=np.uint32)
    vh = rng.integers(0, P, n, dtype=np.uint32)
    tf = rng.integers(0, P, n, dtype=np.uint32)
    return cols, inter, alpha, chal, vh, tf


def col_array(pin, cols):
    """upload the columns, then their device pointers as a device array"""
    ptrs = [pin.dev(c) for c in cols]
    arr = np.array(ptrs, dtype=np.uint64).view(np.uint32)
    return pin.dev(arr)


class Up(Pin):
    def dev(self, a):
        from gpu_util import dev
        b = dev(a)
        self.bufs.append(b)
        return b.ptr


def gpu_full(cols, inter, alpha, chal, vh, tf, b, n):
    pin = Up()
    dcols = col_array(pin, cols)
    ints = [pin.dev(c) for c in inter]
    outs = [empty(n) for _ in range(4)]
    call("vx_cairo_quotient", dcols, *ints, *[o.ptr for o in outs], pin.dev(alpha), pin.dev(vh), pin.dev(tf),
         pin.dev(chal), n, b, None)
    sync()
    return [host(o) for o in outs]


@pytest.mark.parametrize("log_n,b", [(4, 4), (10, 4), (14, 2), (16, 4)])
def test_cairo_quotient_matches_oracle(log_n, b):
    n = 1 << log_n
    cols, inter, alpha, chal, vh, tf = rand_inputs(log_n, n)
    print("cairo quotient n", n, "blowup step", b, flush=True)
    got = gpu_full(cols, inter, alpha, chal, vh, tf, b, n)
    want = po.cairo_quotient(cols, inter, alpha, vh, tf, chal, b)
    for k in range(4):
        assert np.array_equal(got[k], want[k]), k


def test_cairo_quotient_non_power_of_two_rows():
    n, b = 1000, 4
    cols, inter, alpha, chal, vh, tf = rand_inputs(3, n)
    got = gpu_full(cols, inter, alpha, chal, vh, tf, b, n)
    want = po.cairo_quotient(cols, inter, alpha, vh, tf, chal, b)
    for k in range(4):
        assert np.array_equal(got[k], want[k])


def test_cairo_quotient_zero_on_valid_rows():
    rng = random.Random(3)
    n, b = 64, 4
    z, a = tuple(rng.randrange(P) for _ in range(4)), tuple(rng.randrange(P) for _ in range(4))
    chal_q = [z, a, pyref.qm_mul(a, a)] + [tuple(rng.randrange(P) for _ in range(4)) for _ in range(3)]
    rows, inter_rows, last = pc.valid_table(rng, n, b, chal_q)
    cols = [np.array([rows[i][c] for i in range(n)], np.uint32) for c in range(34)]
    inter = [np.array([inter_rows[i][g][k] for i in range(n)], np.uint32) for g in range(8) for k in range(4)]
    tf = np.array([0 if last[i] else rng.randrange(1, P) for i in range(n)], np.uint32)
    vh = np.array([rng.randrange(1, P) for _ in range(n)], np.uint32)
    alpha = np.array([rng.randrange(P) for _ in range(160)], np.uint32)
    chal = np.array([w for q in chal_q for w in q], np.uint32)
    got = gpu_full(cols, inter, alpha, chal, vh, tf, b, n)
    assert all(not np.any(g) for g in got)
    cols[21][9] = (int(cols[21][9]) + 1) % P   # one changed dst
    got = gpu_full(cols, inter, alpha, chal, vh, tf, b, n)
    assert any(np.any(g) for g in got)


def test_cairo_quotient_slab_equals_full():
    n, b, chunk = 1 << 12, 4, 1 << 10
    cols, inter, alpha, chal, vh, tf = rand_inputs(21, n)
    full = po.cairo_quotient(cols, inter, alpha, vh, tf, chal, b)
    pin = Up()
    outs = [empty(n) for _ in range(4)]
    dvh, dtf, dal, dch = pin.dev(vh), pin.dev(tf), pin.dev(alpha), pin.dev(chal)
    for off in range(0, n, chunk):
        idx = [(off + r) % n for r in range(chunk + b)]
        dcols = col_array(pin, [c[idx] for c in cols])
        ints = [pin.dev(c[idx]) for c in inter]
        call("vx_cairo_quotient_slab", dcols, *ints, *[o.ptr for o in outs], dal, dvh, dtf, dch, chunk, b, off, None)
    sync()
    for k in range(4):
        assert np.array_equal(host(outs[k]), full[k])


def test_cairo_quotient_chunked_form():
    n, b = 1 << 10, 4
    cols, inter, alpha, chal, vh, tf = rand_inputs(8, n)
    inter12 = inter[0:4] + inter[16:20] + inter[28:32]
    off, chunk = 256, 512
    want = po.cairo_quotient_chunk(cols, inter12, alpha[:140], vh, tf, chal, off, chunk, b)
    pin = Up()
    dcols = col_array(pin, cols)
    outs = [empty(chunk) for _ in range(4)]
    call("vx_cairo_quotient_chunk", dcols, *[pin.dev(c) for c in inter12], *[o.ptr for o in outs],
         pin.dev(alpha[:140]), pin.dev(vh), pin.dev(tf), pin.dev(chal), off, chunk, n, b, None)
    sync()
    for k in range(4):
        assert np.array_equal(host(outs[k]), want[k])


@pytest.mark.parametrize("log_eval,log_n", [(6, 4), (12, 10), (20, 18)])
def test_vanishing(log_eval, log_n):
    ev = po.eval_domain(log_eval)
    for inverse in (False, True):
        out = empty(1 << log_eval)
        call("vx_compute_vanishing_inv" if inverse else "vx_compute_vanishing", int(ev[0]), int(ev[1]), int(ev[2]),
             int(ev[3]), out.ptr, log_eval, log_n, None)
        sync()
        want = po.cairo_vanishing(ev[:2], ev[2:4], log_eval, log_n, inverse=inverse)
        assert np.array_equal(host(out), want)


def test_axpy():
    rng = np.random.default_rng(2)
    x = rng.integers(0, P, 5000, dtype=np.uint32)
    y = rng.integers(0, P, 5000, dtype=np.uint32)
    pin = Up()
    dy = empty(5000)
    from gpu_util import H2D
    import ctypes as C
    call("vx_memcpy", C.c_void_p(dy.ptr), C.c_void_p(y.ctypes.data), y.nbytes, H2D)
    call("vx_axpy_m31", 777, pin.dev(x), dy.ptr, 5000, None)
    sync()
    assert np.array_equal(host(dy), po.axpy_m31(777, x, y))


def levels_words(log_n, store_lo):
    return sum(8 << (log_n - l) for l in range(store_lo, log_n + 1))


def gpu_commit(cols, inter, alpha, chal, vh, tf, b, log_n, store_lo):
    import ctypes as C
    n = 1 << log_n
    pin = Up()
    dcols = col_array(pin, cols)
    hint = (C.c_void_p * 32)(*[pin.dev(c) for c in inter])
    outs = [empty(n) for _ in range(4)]
    lv, root = empty(levels_words(log_n, store_lo)), empty(8)
    call("vx_cairo_quotient_commit", dcols, hint, *[o.ptr for o in outs], pin.dev(alpha), pin.dev(vh), pin.dev(tf),
         pin.dev(chal), log_n, b, store_lo, lv.ptr, root.ptr, None)
    return [host(o) for o in outs], host(root)


@pytest.mark.parametrize("log_n", [6, 12, 17, 20])
def test_cairo_quotient_commit_matches_oracle(log_n):
    n = 1 << log_n
    cols, inter, alpha, chal, vh, tf = rand_inputs(100 + log_n, n)
    print("cairo commit log_n", log_n, flush=True)
    q, root = gpu_commit(cols, inter, alpha, chal, vh, tf, 4, log_n, min(log_n, 8))
    want = po.cairo_quotient(cols, inter, alpha, vh, tf, chal, 4)
    for k in range(4):
        assert np.array_equal(q[k], want[k]), k
    assert np.array_equal(root, po.merkle_root(want))


@pytest.mark.parametrize("log_n,store_lo,tile_log,lpt",
                         [(13, 4, 12, 4), (13, 3, 11, 3), (13, 2, 10, 2), (14, 2, 8, 0), (12, 5, 12, 4)])
def test_cairo_quotient_commit_every_tile_shape(tmp_path, log_n, store_lo, tile_log, lpt):
    # each subtree shape pinned through the tuner's cache file, for a key
    # (leaf kind 7, log size, stored levels) that no other test commits; the
    # shapes with 4..16 leaves per lane compute the tile's rows first, one
    # per lane, and hash them after the workgroup's barrier
    import vortexstark_amd as vx
    f = tmp_path / "tune.txt"
    f.write_text(f"7 0 0 0 {log_n} {store_lo} {tile_log} {lpt}\n")
    vx.tuner(True, str(f))
    try:
        ents = vx.tuner_entries()
        assert any((e["leaf_kind"], e["log_n"], e["tile_log"], e["lane_log"]) == (7, log_n, tile_log, lpt)
                   for e in ents), ents
        n = 1 << log_n
        cols, inter, alpha, chal, vh, tf = rand_inputs(300 + tile_log + lpt, n)
        print("cairo commit shape", log_n, store_lo, tile_log, lpt, flush=True)
        q, root = gpu_commit(cols, inter, alpha, chal, vh, tf, 4, log_n, store_lo)
        want = po.cairo_quotient(cols, inter, alpha, vh, tf, chal, 4)
        for k in range(4):
            assert np.array_equal(q[k], want[k]), k
        assert np.array_equal(root, po.merkle_root(want))
    finally:
        vx.tuner(True, None)


def test_cairo_shaped_commitment_matches_oracle_digest(golden):
    # 34 trace columns committed as lo / hi / dict trees (16, 15, 3 columns,
    # src/cairo_air/prover.rs:341-350) and the fused quotient commitment,
    # inputs from the synthetic generator (vx_m31_fill_random)
    import ctypes as C
    g = golden["cairo_commit_c"]
    L, seed, b = g["log_eval"], g["seed"], g["blowup_step"]
    n = 1 << L
    bufs = [empty(n) for _ in range(70)]
    for c, buf in enumerate(bufs):
        call("vx_m31_fill_random", buf.ptr, L, seed, c, None)
    small = [empty(1 << 8), empty(1 << 5)]
    call("vx_m31_fill_random", small[0].ptr, 8, seed, 68, None)
    call("vx_m31_fill_random", small[1].ptr, 5, seed, 69, None)
    print("cairo-shaped commitment log_eval", L, flush=True)
    roots = []
    for lo, hi in ((0, 16), (16, 31), (31, 34)):
        lv, root = empty(levels_words(L, 10)), empty(8)
        hc = (C.c_void_p * (hi - lo))(*[bufs[c].ptr for c in range(lo, hi)])
        call("vx_merkle_commit_tiled", hc, hi - lo, L, 10, lv.ptr, root.ptr, None)
        roots.append([int(x) for x in host(root)])
    pin = Up()
    dcols = pin.dev(np.array([bufs[c].ptr for c in range(34)], np.uint64).view(np.uint32))
    hint = (C.c_void_p * 32)(*[bufs[34 + k].ptr for k in range(32)])
    outs = [empty(n) for _ in range(4)]
    lv, root = empty(levels_words(L, 10)), empty(8)
    call("vx_cairo_quotient_commit", dcols, hint, *[o.ptr for o in outs], small[0].ptr, bufs[66].ptr, bufs[67].ptr,
         small[1].ptr, L, b, 10, lv.ptr, root.ptr, None)
    roots.append([int(x) for x in host(root)])
    assert roots == g["roots"]
