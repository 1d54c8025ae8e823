"""Batched uneven columns on the GPU: the work-list LDE, the batched bit
reverse and the lifted Merkle tree against the oracle, column by column and
root by root; the resident batch against the committed workload digest."""
import ctypes as C
import hashlib

import numpy as np
import pytest

import pyoracle as po
import vortexstark_amd as vx
from gpu_util import d2h, dev, empty, host, sync

pytestmark = pytest.mark.gpu


def ptrs(bufs):
    return (C.c_void_p * len(bufs))(*[b.ptr for b in bufs])


def u32s(xs):
    return (C.c_uint32 * len(xs))(*xs)


def bitrev(a):
    n = a.size
    bits = n.bit_length() - 1
    idx = np.arange(n, dtype=np.uint64)
    r = np.zeros(n, dtype=np.uint64)
    for b in range(bits):
        r |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(bits - 1 - b)
    out = np.empty_like(a)
    out[r.astype(np.int64)] = a
    return out


def test_lde_batch_matches_oracle_per_column():
    rng = np.random.default_rng(11)
    logs = [14, 16, 14, 15, 17, 14]   # uneven, unsorted on purpose (no ordering needed here)
    bb = 2
    vals = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in logs]
    ins = [dev(v) for v in vals]
    coefs = [empty(1 << L) for L in logs]
    evals = [empty(1 << (L + bb)) for L in logs]
    vx._lib.call("vx_circle_ntt_lde_batch", ptrs(ins), ptrs(coefs), ptrs(evals), u32s(logs), len(logs), bb, None)
    for i, L in enumerate(logs):
        print("column", i, L, flush=True)
        want_c = po.interpolate(vals[i], po.coset(po.HALF_COSET, L))
        assert np.array_equal(host(coefs[i]), want_c)
        assert np.array_equal(host(evals[i]), po.lde(vals[i], bb))
        assert np.array_equal(host(ins[i]), vals[i])  # inputs kept intact


def test_bit_reverse_batch():
    rng = np.random.default_rng(12)
    logs = [10, 13, 11, 16, 14, 15, 12]
    vals = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in logs]
    src = [dev(v) for v in vals]
    dst = [empty(1 << L) for L in logs]
    vx._lib.call("vx_bit_reverse_m31_batch", ptrs(src), ptrs(dst), u32s(logs), len(logs), None)
    for v, d in zip(vals, dst):
        assert np.array_equal(host(d), bitrev(v))


def test_bit_reverse_batch_unaligned_columns():
    # columns off 16-byte alignment take the 4-byte path, aligned ones in the
    # same list the 16-byte path
    rng = np.random.default_rng(13)
    logs = [12, 10, 14]
    vals = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in logs]
    pad_src, off = [1, 0, 1], [1, 0, 0]  # words before each column in its buffer
    src = [dev(np.concatenate([[1] * p, v]).astype(np.uint32)) for p, v in zip(pad_src, vals)]
    dst = [empty((1 << L) + 1) for L in logs]
    sp = (C.c_void_p * 3)(*[b.ptr + 4 * p for b, p in zip(src, pad_src)])
    dp = (C.c_void_p * 3)(*[b.ptr + 4 * o for b, o in zip(dst, off)])
    vx._lib.call("vx_bit_reverse_m31_batch", sp, dp, u32s(logs), len(logs), None)
    for v, d, o in zip(vals, dst, off):
        assert np.array_equal(host(d)[o:o + v.size], bitrev(v))


@pytest.mark.parametrize("logs", [[4], [3, 3, 4, 6], [2] * 7 + [4] * 9 + [6] * 16, [5, 6, 8, 8, 9, 12, 12, 13]])
def test_lifted_leaves_and_root(logs):
    rng = np.random.default_rng(len(logs))
    cols = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in logs]
    bufs = [dev(c) for c in cols]
    S = logs[-1]
    out = empty(8 << S)
    vx._lib.call("vx_build_leaves_lifted", ptrs(bufs), u32s(logs), len(logs), S, out.ptr, None)
    assert np.array_equal(host(out).reshape(-1, 8), po.leaves_lifted(cols, S))
    if S >= 4:
        root = empty(8)
        vx._lib.call("vx_merkle_root_lifted", ptrs(bufs), u32s(logs), len(logs), root.ptr, None)
        assert np.array_equal(host(root), po.merkle_root_lifted(cols))


def test_resident_batch_matches_oracle(golden):
    g = next(x for x in golden["batch_c"] if x["n_cols"] == 12)
    b = vx.Batch(g["log_sizes"], g["blowup_bits"], g["bit_reverse"])
    b.fill_random(g["seed"])
    for c, L in enumerate(g["log_sizes"][:3]):
        ptr, words = b.column(c, 0)
        assert np.array_equal(d2h(ptr, words), po.d2_column(g["seed"], c, L))
    root = b.run()
    assert root == g["root"]
    # every output column equals the oracle's bit-reversed LDE
    for c in (0, 5, 11):
        L = g["log_sizes"][c]
        ptr, words = b.column(c, 2)
        e = po.lde(po.d2_column(g["seed"], c, L), g["blowup_bits"])
        po.lib().or_bit_reverse_m31(e, L + g["blowup_bits"])
        assert np.array_equal(d2h(ptr, words), e)
    # a second run gives the same root (inputs are not consumed)
    assert b.run() == root
    b.close()


def test_batched_workload_digest(golden):
    g = next((x for x in golden["batch_c"] if x["n_cols"] == 96), None)
    if g is None:
        pytest.skip("96-column digest not in golden.json (gen_golden.py --big)")
    b = vx.Batch(g["log_sizes"], g["blowup_bits"], g["bit_reverse"])
    b.fill_random(g["seed"])
    assert b.run(timed=True) == g["root"]
    t = b.stage_times()
    print("batch stage ms", t, flush=True)
    assert t["total"] > 0
    b.close()


def test_batch_rejects_bad_shapes():
    with pytest.raises(vx.VortexError, match="ascending"):
        vx.Batch([15, 14])
    with pytest.raises(vx.VortexError, match="14..24"):
        vx.Batch([12, 14])
    sync()


@pytest.mark.parametrize("logs,bb,pw", [([14], 1, 0), ([14, 14, 15, 16], 2, 0), ([14] * 17 + [15], 2, 8),
                                        ([15, 15, 16, 17, 17, 17], 1, 0)])
def test_batch_proofs_match_oracle(logs, bb, pw):
    print("batch proof", logs, bb, pw, flush=True)
    p = vx.BatchProver(logs, bb, 80, pw)
    for seed in (7, 2**40 + 3):
        got = p.prove(seed)
        want = po.batch_prove(seed, logs, bb, 80, pw)
        assert len(got) == len(want) == p.size
        assert got == want, (logs, seed)
    assert vx.verify(got, bb) == (True, "")
    p.close()


def test_batch_proof_graph_replay_equals_eager():
    logs = [14, 15, 15, 16]
    g = vx.BatchProver(logs, graphs=True)
    e = vx.BatchProver(logs, graphs=False)
    for seed in (1, 2, 1, 99):
        assert g.prove(seed) == e.prove(seed)
    g.close()
    e.close()


@pytest.mark.parametrize("n_cols,pw", [(12, 0), (96, 0), (96, 20)])
def test_batch_proof_workload_digests(golden, n_cols, pw):
    g = next((x for x in golden["proofs_batch_c"] if (x["n_cols"], x["pow_bits"]) == (n_cols, pw)), None)
    if g is None:
        pytest.skip("digest not in golden.json (gen_golden.py --big)")
    p = vx.BatchProver(g["log_sizes"], g["blowup_bits"], g["n_queries"], pw)
    got = p.prove(g["seed"])
    assert len(got) == g["len"]
    assert hashlib.sha256(got).hexdigest() == g["sha256"]
    assert vx.verify(got) == (True, "")
    p.close()
