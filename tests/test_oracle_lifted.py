"""Batched uneven columns on the CPU: the oracle's per-column LDE against the
pinned interpolate/evaluate, and the lifted Merkle tree
(cuda/merkle_leaves_lifted.cu) against a hashlib restatement."""
import hashlib
import struct

import numpy as np

import pyoracle as po
import pyref


def test_lde_is_interpolate_then_evaluate():
    rng = np.random.default_rng(1)
    for log_n, bb in [(4, 1), (6, 2), (8, 3)]:
        v = rng.integers(0, po.P, size=1 << log_n, dtype=np.uint32)
        coef = po.interpolate(v, po.coset(po.HALF_COSET, log_n))
        pad = np.zeros(1 << (log_n + bb), np.uint32)
        pad[:1 << log_n] = coef
        assert np.array_equal(po.lde(v, bb), po.evaluate(pad, po.eval_domain(log_n + bb)))


def lifted_row(i, S, L):
    return i if L == S else ((i >> (S - L + 1)) << 1) | (i & 1)


def hashlib_leaves(cols, S):
    logs = [len(c).bit_length() - 1 for c in cols]
    out = []
    for i in range(1 << S):
        data = b"".join(struct.pack("<I", int(c[lifted_row(i, S, L)])) for c, L in zip(cols, logs))
        out.append(list(struct.unpack("<8I", hashlib.blake2s(data).digest())))
    return np.array(out, np.uint32)


def test_lifted_leaves_match_hashlib():
    rng = np.random.default_rng(2)
    # 1..40 columns (one, two and three Blake2s blocks, a full last block)
    for sizes in [[5], [3, 3, 4, 6], [2] * 7 + [4] * 9 + [6] * 16, [1, 2, 3, 4, 5, 6] * 6 + [6] * 4]:
        sizes = sorted(sizes)
        cols = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in sizes]
        S = sizes[-1]
        assert np.array_equal(po.leaves_lifted(cols, S), hashlib_leaves(cols, S)), sizes
    # lifting above the largest column
    cols = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in (2, 3)]
    assert np.array_equal(po.leaves_lifted(cols, 5), hashlib_leaves(cols, 5))


def test_lifted_row_rule():
    # a column of the tree's size is read in order; a half-size column
    # repeats each pair (2j, 2j+1) over leaves (4j, 4j+1) and (4j+2, 4j+3)
    assert [lifted_row(i, 4, 4) for i in range(16)] == list(range(16))
    assert [lifted_row(i, 4, 3) for i in range(8)] == [0, 1, 0, 1, 2, 3, 2, 3]
    assert [lifted_row(i, 4, 2) for i in range(8)] == [0, 1, 0, 1, 0, 1, 0, 1]


def test_lifted_root_is_plain_tree_over_lifted_leaves():
    rng = np.random.default_rng(3)
    sizes = [3, 4, 4, 5, 6]
    cols = [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in sizes]
    leaves = hashlib_leaves(cols, 6)
    layers = pyref.merkle_layers_from_hashes([bytes(np.asarray(h, "<u4").tobytes()) for h in leaves])
    assert list(po.merkle_root_lifted(cols)) == list(struct.unpack("<8I", layers[-1][0]))
