"""The Cairo VM AIR quotient in the C oracle (or_cairo_quotient,
or_cairo_quotient_chunk; cuda/cairo_constraint.cu) against the independent
Python restatement (tests/golden/pyref_cairo.py, written from the reference's
host-side evaluation, src/cairo_air/prover.rs:4294-4580): random rows in the
full, slab and chunk forms; valid Cairo tables give a zero quotient wherever
the transition is enforced; one changed value makes its row nonzero. Also the
vanishing polynomial helper and axpy."""
import random

import numpy as np
import pytest

import pyoracle as po
import pyref
import pyref_cairo as pc

P = po.P


def rand_q(rng):
    return tuple(rng.randrange(P) for _ in range(4))


def challenges(rng):
    z, a = rand_q(rng), rand_q(rng)
    return [z, a, pyref.qm_mul(a, a), rand_q(rng), rand_q(rng), rand_q(rng)]


def flat(qs):
    return np.array([w for q in qs for w in q], np.uint32)


def soa(rows_q):
    """list over rows of QM31 tuples -> 4 numpy columns"""
    return [np.array([q[k] for q in rows_q], np.uint32) for k in range(4)]


def random_inputs(rng, n, binary_flags=True):
    cols = []
    for c in range(34):
        if binary_flags and (5 <= c < 20 or c == 33):
            cols.append([rng.randrange(2) for _ in range(n)])
        else:
            cols.append([rng.randrange(P) for _ in range(n)])
    inter = [[rand_q(rng) for _ in range(n)] for _ in range(8)]
    return cols, inter


def oracle_full(cols, inter, alphas, vh, tf, chal, b):
    inter_cols = [c for it in inter for c in soa(it)]
    return po.cairo_quotient([np.array(c, np.uint32) for c in cols], inter_cols, flat(alphas),
                             np.array(vh, np.uint32), np.array(tf, np.uint32), flat(chal), b)


def python_full(cols, inter, alphas, vh, tf, chal, b):
    n = len(cols[0])
    out = []
    for i in range(n):
        j = (i + b) % n
        row = [cols[c][i] for c in range(34)]
        nxt = [cols[c][j] for c in range(34)]
        out.append(pc.quotient_row(row, nxt, [it[i] for it in inter], [it[j] for it in inter], alphas, chal,
                                   tf[i], vh[i]))
    return out


@pytest.mark.parametrize("n,b,binary", [(16, 4, True), (8, 2, False), (32, 4, True)])
def test_full_quotient_matches_python(n, b, binary):
    rng = random.Random(n * 7 + b)
    cols, inter = random_inputs(rng, n, binary)
    alphas = [rand_q(rng) for _ in range(40)]
    chal = challenges(rng)
    vh = [rng.randrange(P) for _ in range(n)]
    tf = [rng.randrange(P) for _ in range(n)]
    got = oracle_full(cols, inter, alphas, vh, tf, chal, b)
    want = python_full(cols, inter, alphas, vh, tf, chal, b)
    for k in range(4):
        assert [int(x) for x in got[k]] == [q[k] for q in want]


def test_slab_equals_full():
    # slabs of chunk + blowup_step rows (the last one wraps to the first
    # rows) give the full-mode rows of their range
    rng = random.Random(5)
    n, b, chunk = 32, 4, 8
    cols, inter = random_inputs(rng, n)
    alphas = [rand_q(rng) for _ in range(40)]
    chal = challenges(rng)
    vh = np.array([rng.randrange(P) for _ in range(n)], np.uint32)
    tf = np.array([rng.randrange(P) for _ in range(n)], np.uint32)
    full = oracle_full(cols, inter, alphas, vh, tf, chal, b)
    inter_cols = [np.array(c, np.uint32) for it in inter for c in soa(it)]
    cols_np = [np.array(c, np.uint32) for c in cols]
    for off in range(0, n, chunk):
        idx = [(off + r) % n for r in range(chunk + b)]
        got = po.cairo_quotient([c[idx] for c in cols_np], [c[idx] for c in inter_cols], flat(alphas), vh, tf,
                                flat(chal), b, n=chunk, chunk_offset=off, slab=True)
        for k in range(4):
            assert np.array_equal(got[k][off:off + chunk], full[k][off:off + chunk])


def test_chunk_quotient_matches_python():
    rng = random.Random(11)
    n, b = 16, 4
    cols, inter = random_inputs(rng, n)
    s3 = [inter[0], inter[4], inter[7]]
    alphas = [rand_q(rng) for _ in range(35)]
    chal = challenges(rng)
    vh = [rng.randrange(P) for _ in range(n)]
    tf = [rng.randrange(P) for _ in range(n)]
    off, chunk = 4, 8
    got = po.cairo_quotient_chunk([np.array(c, np.uint32) for c in cols], [c for it in s3 for c in soa(it)],
                                  flat(alphas), np.array(vh, np.uint32), np.array(tf, np.uint32), flat(chal), off,
                                  chunk, b)
    for r in range(chunk):
        i = off + r
        j = (i + b) % n
        want = pc.quotient_row_chunk([cols[c][i] for c in range(34)], [cols[c][j] for c in range(34)],
                                     [s[i] for s in s3], [s[j] for s in s3], alphas, chal, tf[i], vh[i])
        assert tuple(int(got[k][r]) for k in range(4)) == want, r


def valid_inputs(seed, n, b):
    rng = random.Random(seed)
    chal = challenges(rng)
    rows, inter, last = pc.valid_table(rng, n, b, chal)
    cols = [[rows[i][c] for i in range(n)] for c in range(34)]
    inter_t = [[inter[i][k] for i in range(n)] for k in range(8)]
    tf = [0 if last[i] else rng.randrange(1, P) for i in range(n)]
    vh = [rng.randrange(1, P) for _ in range(n)]
    return rng, cols, inter_t, chal, tf, vh


@pytest.mark.parametrize("seed,n,b", [(1, 16, 4), (2, 32, 2), (3, 64, 4), (4, 8, 8)])
def test_valid_cairo_rows_give_zero(seed, n, b):
    rng, cols, inter, chal, tf, vh = valid_inputs(seed, n, b)
    alphas = [rand_q(rng) for _ in range(40)]
    got = oracle_full(cols, inter, alphas, vh, tf, chal, b)
    assert all(not np.any(g) for g in got)
    # the chunked form (35 constraints, inverse sums) agrees on valid rows
    s3 = [inter[0], inter[4], inter[7]]
    gc = po.cairo_quotient_chunk([np.array(c, np.uint32) for c in cols], [c for it in s3 for c in soa(it)],
                                 flat([rand_q(rng) for _ in range(35)]), np.array(vh, np.uint32),
                                 np.array(tf, np.uint32), flat(chal), 0, n, b)
    assert all(not np.any(g) for g in gc)


@pytest.mark.parametrize("col", [0, 1, 2, 3, 7, 12, 20, 21, 25, 26, 27, 31, 32, 33])
def test_one_changed_value_is_caught(col):
    n, b = 16, 4
    rng, cols, inter, chal, tf, vh = valid_inputs(100 + col, n, b)
    alphas = [rand_q(rng) for _ in range(40)]
    # a row whose transition and whose predecessor's transition are both
    # enforced; the dict key and value enter only on active dict rows
    cand = [r for r in range(n) if tf[r] and tf[(r - b) % n] and (col not in (31, 32) or cols[33][r] == 1)]
    assert cand
    i = cand[0]
    # the changed row's own constraints, or the previous row's transition
    # (next-row registers), must see the change; flags stay binary
    cols[col][i] = 1 - cols[col][i] if 5 <= col < 20 or col == 33 else (cols[col][i] + 12345) % P
    got = oracle_full(cols, inter, alphas, vh, tf, chal, b)
    touched = {i, (i - b) % n}
    bad = {r for r in range(n) if any(int(g[r]) for g in got)}
    assert bad & touched, (col, bad)


def test_dst_inv_is_checked_on_jumping_jnz_rows():
    # dst_inv only enters where jnz = 1 and dst != 0 (constraints 23, 24)
    n, b = 32, 4
    rng, cols, inter, chal, tf, vh = valid_inputs(77, n, b)
    alphas = [rand_q(rng) for _ in range(40)]
    rows = [i for i in range(n) if cols[14][i] == 1 and cols[21][i] != 0]
    assert rows
    i = rows[0]
    cols[30][i] = (cols[30][i] + 1) % P
    got = oracle_full(cols, inter, alphas, vh, tf, chal, b)
    assert any(int(g[i]) for g in got)


def test_interaction_change_is_caught():
    n, b = 16, 4
    rng, cols, inter, chal, tf, vh = valid_inputs(7, n, b)
    alphas = [rand_q(rng) for _ in range(40)]
    for k in range(8):
        it = [list(col) for col in inter]
        it[k][9] = pyref.qm_add(it[k][9], (1, 0, 0, 0))
        got = oracle_full(cols, it, alphas, vh, tf, chal, b)
        assert any(int(g[9]) or int(g[(9 - b) % n]) for g in got), k


def test_vanishing_zero_on_trace_domain_and_inverse():
    for log_n in (2, 4, 7):
        tr = po.coset(po.HALF_COSET, log_n)
        z = po.cairo_vanishing(tr[:2], tr[2:4], log_n, log_n, inverse=False)
        assert not np.any(z)
        ev = po.eval_domain(log_n + 2)
        zs = po.cairo_vanishing(ev[:2], ev[2:4], log_n + 2, log_n, inverse=False)
        zi = po.cairo_vanishing(ev[:2], ev[2:4], log_n + 2, log_n, inverse=True)
        assert np.all(zs != 0)
        assert all(int(a) * int(b) % P == 1 for a, b in zip(zs, zi))
        # position i holds the point at bit-reversed index i (pyref geometry)
        c = pyref.eval_domain(log_n + 2)
        for i in (0, 1, 5, (1 << (log_n + 2)) - 1):
            x = pyref.coset_at(c, pyref.bitrev(i, log_n + 2))[0]
            for _ in range(log_n):
                x = (2 * x * x - 1) % P
            assert int(zs[i]) == (x + 1) % P


def test_axpy():
    rng = np.random.default_rng(1)
    x = rng.integers(0, P, 100, dtype=np.uint32)
    y = rng.integers(0, P, 100, dtype=np.uint32)
    got = po.axpy_m31(12345, x, y)
    assert [int(v) for v in got] == [(int(b) + 12345 * int(a)) % P for a, b in zip(x, y)]
