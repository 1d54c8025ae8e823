"""Pins the oracle's stwo-basis circle NTT (oracle/vx_oracle.c, or_stwo_*).

The reference holds no vectors for this path (its own checks compare with the
stwo crate, which is not vendored), so the restatement is pinned by what the
mathematics fixes:
  * evaluation at a point against the basis definition, term by term
    (1, y, x, xy, 2x^2 - 1, ...; src/oods.rs:133-161), with an independent
    pure-Python QM31;
  * the FFT against that direct sum at every point of the CircleDomain in
    bit-reversed order, for full random coefficient vectors (every layer's
    twiddles matter), which is also the reference's own check
    (src/ntt.rs:517-548, coefficients 42, 17, 5, 99 at log_n 8);
  * interpolation of the coordinate functions x and y, the roundtrip
    (src/ntt.rs:923-935) and the twiddle/inverse-twiddle pairing.
"""
import random

import numpy as np
import pytest

import pyoracle as po
import pyref

P = po.P


def bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def m31_point(x, y):
    return (x, 0, 0, 0), (y, 0, 0, 0)


def basis_value(idx, zx, zy):
    """prod over the set bits b of idx of m_b: m_0 = y, m_1 = x, m_b = 2 m_{b-1}^2 - 1."""
    r, m, b = (1, 0, 0, 0), zx, 0
    while idx >> b:
        f = zy if b == 0 else m
        if b >= 2:
            m2 = pyref.qm_mul(m, m)
            m = pyref.qm_sub(pyref.qm_add(m2, m2), (1, 0, 0, 0))
            f = m
        if (idx >> b) & 1:
            r = pyref.qm_mul(r, f)
        b += 1
    return r


def test_eval_at_point_closed_forms():
    rng = random.Random(3)
    zx = tuple(rng.randrange(P) for _ in range(4))
    zy = tuple(rng.randrange(P) for _ in range(4))
    x2 = pyref.qm_mul(zx, zx)
    dx = pyref.qm_sub(pyref.qm_add(x2, x2), (1, 0, 0, 0))  # 2x^2 - 1
    ddx = pyref.qm_sub(pyref.qm_add(pyref.qm_mul(dx, dx), pyref.qm_mul(dx, dx)), (1, 0, 0, 0))
    expect = {0: (1, 0, 0, 0), 1: zy, 2: zx, 3: pyref.qm_mul(zx, zy), 4: dx, 5: pyref.qm_mul(dx, zy),
              6: pyref.qm_mul(dx, zx), 8: ddx, 15: pyref.qm_mul(pyref.qm_mul(ddx, dx), pyref.qm_mul(zx, zy))}
    for idx, v in expect.items():
        c = np.zeros(16, np.uint32)
        c[idx] = 1
        assert po.stwo_eval_at_point(c, zx, zy) == v, idx
    # linear in the coefficients, and the general term
    c = np.array([rng.randrange(P) for _ in range(32)], np.uint32)
    acc = (0, 0, 0, 0)
    for i in range(32):
        acc = pyref.qm_add(acc, pyref.qm_scale(basis_value(i, zx, zy), int(c[i])))
    assert po.stwo_eval_at_point(c, zx, zy) == acc


def test_twiddle_buffer_shape():
    for log_n in range(2, 12):
        fwd, inv = po.stwo_twiddles(log_n)
        assert fwd.size == 1 << (log_n - 1) and fwd[-1] == 1
        assert all(int(a) * int(b) % P == 1 for a, b in zip(fwd, inv))
        assert np.all(fwd < P) and np.all(fwd > 0)


def test_reference_vector_eval_domain_log8():
    # src/ntt.rs:517-548: value k is f at CircleDomain point bitrev(k)
    log_n = 8
    c = np.zeros(1 << log_n, np.uint32)
    c[:4] = [42, 17, 5, 99]
    ev = po.stwo_evaluate(c)
    for k in range(1 << log_n):
        zx, zy = m31_point(*po.stwo_domain_point(log_n, bitrev(k, log_n)))
        assert po.stwo_eval_at_point(c, zx, zy) == (int(ev[k]), 0, 0, 0), k


@pytest.mark.parametrize("log_n", [3, 4, 5, 7, 9])
def test_evaluate_is_the_direct_sum_at_every_point(log_n):
    rng = np.random.default_rng(log_n)
    c = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    ev = po.stwo_evaluate(c)
    for k in range(1 << log_n):
        zx, zy = m31_point(*po.stwo_domain_point(log_n, bitrev(k, log_n)))
        assert po.stwo_eval_at_point(c, zx, zy) == (int(ev[k]), 0, 0, 0), (log_n, k)


def test_four_points_use_the_pad_word():
    """At n = 4 the twiddle buffer is [x0, 1]: the first line layer has one
    word, so the circle layer's (x, y) pair is (x0, pad 1) and its twiddles are
    (1, -1, -x0, x0) (cuda/circle_ntt_stwo.cu:214-234). The reference's values
    there are this closed form, not the CircleDomain evaluation."""
    fwd, _ = po.stwo_twiddles(2)
    x0 = int(fwd[0])
    assert int(fwd[1]) == 1
    rng = random.Random(5)
    for _ in range(20):
        c = [rng.randrange(P) for _ in range(4)]
        a0, a2 = (c[0] + x0 * c[2]) % P, (c[0] - x0 * c[2]) % P  # line layer, twiddle x0
        a1, a3 = (c[1] + x0 * c[3]) % P, (c[1] - x0 * c[3]) % P
        want = [(a0 + a1) % P, (a0 - a1) % P, (a2 - a3) % P, (a2 + a3) % P]  # circle twiddles 1, -1
        assert po.stwo_evaluate(np.array(c, np.uint32)).tolist() == want


@pytest.mark.parametrize("log_n", [3, 4, 6, 10])
def test_interpolate_coordinate_functions(log_n):
    n = 1 << log_n
    pts = [po.stwo_domain_point(log_n, bitrev(k, log_n)) for k in range(n)]
    for f, idx in ((lambda p: 1, 0), (lambda p: p[1], 1), (lambda p: p[0], 2),
                   (lambda p: p[0] * p[1] % P, 3), (lambda p: (2 * p[0] * p[0] - 1) % P, 4)):
        coef = po.stwo_interpolate(np.array([f(p) for p in pts], np.uint32))
        want = np.zeros(n, np.uint32)
        want[idx] = 1
        assert np.array_equal(coef, want), (log_n, idx)


def test_roundtrip():
    # src/ntt.rs:923-935
    d = np.array([(i * 7 + 3) % P for i in range(64)], np.uint32)
    assert np.array_equal(po.stwo_interpolate(po.stwo_evaluate(d)), d)
    rng = np.random.default_rng(11)
    for log_n in (2, 3, 8, 13, 15):
        d = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
        assert np.array_equal(po.stwo_evaluate(po.stwo_interpolate(d)), d)
        assert np.array_equal(po.stwo_interpolate(po.stwo_evaluate(d)), d)


def test_differs_from_the_fibonacci_basis():
    # the stwo basis is not the Fibonacci prover's (src/ntt.rs:626-631)
    c = np.arange(1, 65, dtype=np.uint32)
    assert not np.array_equal(po.stwo_evaluate(c), po.evaluate(c, po.coset(po.HALF_COSET, 6)))


# ---- evaluation from folding factors (cuda_eval_at_point) and barycentric sums

def test_factor_fold_equals_point_evaluation():
    # two restatements: halves folded with factors[0] first (circle_ntt.cu:
    # 429-468) against adjacent pairs folded from z.y up (src/oods.rs)
    rng = np.random.default_rng(21)
    for log_n in range(0, 11):
        zx = tuple(int(v) for v in rng.integers(0, P, 4))
        zy = tuple(int(v) for v in rng.integers(0, P, 4))
        c = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
        f = po.folding_factors(zx, zy, log_n)
        assert po.eval_at_point_factors(c, f) == po.stwo_eval_at_point(c, zx, zy), log_n
    # factors[0] folds the top coefficient bit: a unit vector at the top index
    # evaluates to factors[0]
    f = np.arange(1, 4 * 6 + 1, dtype=np.uint32)
    c = np.zeros(64, np.uint32)
    c[32] = 1
    assert po.eval_at_point_factors(c, f) == (1, 2, 3, 4)
    c[:] = 0
    c[1] = 1
    assert po.eval_at_point_factors(c, f) == tuple(range(21, 25))


def test_barycentric_partials():
    rng = np.random.default_rng(22)
    for n, nb in ((1, 1), (300, 2), (1000, 3), (5000, 7)):
        e = rng.integers(0, P, n, dtype=np.uint32)
        w = rng.integers(0, P, 4 * n, dtype=np.uint32)
        parts = po.barycentric_partials(e, w, nb).reshape(nb, 4).astype(object)
        total = [sum(int(e[i]) * int(w[4 * i + k]) for i in range(n)) % P for k in range(4)]
        assert [int(v) % P for v in parts.sum(axis=0) % P] == total
        # the partition: element i lands in block (i // 256) % n_blocks
        for i in (0, n - 1, n // 2):
            u = np.zeros(n, np.uint32)
            u[i] = 1
            p1 = po.barycentric_partials(u, w, nb).reshape(nb, 4)
            b = (i // 256) % nb
            assert p1[b].tolist() == w[4 * i:4 * i + 4].tolist()
            assert not np.any(np.delete(p1, b, axis=0))


# ---- QuotientOps (cuda/quotient.cu)

def test_accumulate_numerators_closed_forms():
    rng = np.random.default_rng(41)
    n = 64
    cols = [rng.integers(0, P, n, dtype=np.uint32) for _ in range(3)]
    b = rng.integers(0, P, 12, dtype=np.uint32)
    zero = [np.zeros(n, np.uint32)] * 3
    out = po.accumulate_numerators(zero, [0, 1, 2], b, rng.integers(0, P, 12, dtype=np.uint32), n)
    for k in range(4):  # zero columns: -sum b
        assert np.all(out[k] == (-sum(int(b[4 * i + k]) for i in range(3))) % P)
    one = np.tile(np.array([1, 0, 0, 0], np.uint32), 3)
    out = po.accumulate_numerators(cols, [2, 0, 2], np.zeros(12, np.uint32), one, n)
    want = (2 * cols[2].astype(np.int64) + cols[0]) % P
    assert np.array_equal(out[0], want) and not any(np.any(out[k]) for k in (1, 2, 3))
    u = np.tile(np.array([0, 0, 1, 0], np.uint32), 2)  # c = u moves the sum to word 2
    out = po.accumulate_numerators(cols, [1, 1], np.zeros(8, np.uint32), u, n)
    assert np.array_equal(out[2], (2 * cols[1].astype(np.int64)) % P) and not np.any(out[0])


def _cm_mul(x, y):
    return pyref.cm_mul(x, y)


def test_quotients_combine_inverts_the_denominator():
    rng = np.random.default_rng(42)
    L, n = 6, 64
    for lift_lr in (0, 1, 3):  # accumulator smaller than the lifting domain by 2^lr
        acc_log = L - lift_lr
        sx = [int(v) for v in rng.integers(0, P, 4)]
        sy = [int(v) for v in rng.integers(0, P, 4)]
        fla = [int(v) for v in rng.integers(0, P, 4)]
        numer = [rng.integers(0, P, 1 << acc_log, dtype=np.uint32) for _ in range(4)]
        dxs = rng.integers(0, P, n, dtype=np.uint32)
        dys = rng.integers(0, P, n, dtype=np.uint32)
        out = po.quotients_combine(sx, sy, fla, [numer], [acc_log], dxs, dys, L)
        for row in range(n):
            dx, dy = int(dxs[row]), int(dys[row])
            li = ((row >> (lift_lr + 1)) << 1) + (row & 1)
            full = [(int(numer[k][li]) - fla[k] * dy) % P for k in range(4)]
            # the line denominator through the sample point and its conjugate
            den = pyref.qm_sub(_cm_mul(((sx[0] - dx) % P, sx[1]), (sy[2], sy[3])) + (0, 0),
                               _cm_mul(((sy[0] - dy) % P, sy[1]), (sx[2], sx[3])) + (0, 0))[:2]
            q = [int(out[k][row]) for k in range(4)]
            assert _cm_mul((q[0], q[1]), den) == (full[0], full[1]), (lift_lr, row)
            assert _cm_mul((q[2], q[3]), den) == (full[2], full[3]), (lift_lr, row)


def test_quotients_combine_special_point_and_sum_over_samples():
    rng = np.random.default_rng(43)
    L, n = 5, 32
    dxs = rng.integers(0, P, n, dtype=np.uint32)
    dys = rng.integers(0, P, n, dtype=np.uint32)
    # Im(p.x) = 0 and Im(p.y) = 1: the denominator is Re(p.x) - d.x
    sx = [int(v) for v in rng.integers(0, P, 2)] + [0, 0]
    sy = [int(v) for v in rng.integers(0, P, 2)] + [1, 0]
    numer = [rng.integers(0, P, n, dtype=np.uint32) for _ in range(4)]
    out = po.quotients_combine(sx, sy, [0, 0, 0, 0], [numer], [L], dxs, dys, L)
    for row in range(n):
        den = ((sx[0] - int(dxs[row])) % P, sx[1])
        for h in (0, 2):
            assert _cm_mul((int(out[h][row]), int(out[h + 1][row])), den) == (int(numer[h][row]), int(numer[h + 1][row]))
    # two accumulators give the sum of the two single-accumulator quotients
    sx2 = [int(v) for v in rng.integers(0, P, 4)]
    sy2 = [int(v) for v in rng.integers(0, P, 4)]
    numer2 = [rng.integers(0, P, n // 2, dtype=np.uint32) for _ in range(4)]
    a = po.quotients_combine(sx, sy, [1, 2, 3, 4], [numer], [L], dxs, dys, L)
    b = po.quotients_combine(sx2, sy2, [5, 6, 7, 8], [numer2], [L - 1], dxs, dys, L)
    ab = po.quotients_combine(sx + sx2, sy + sy2, [1, 2, 3, 4, 5, 6, 7, 8], [numer, numer2], [L, L - 1], dxs, dys, L)
    for k in range(4):
        assert np.array_equal(ab[k], (a[k].astype(np.int64) + b[k]) % P)


# ---- GrindOps nonce family (grind_ops.rs:71-128)

def _pow_ok(prefix_words, bits, nonce):
    import hashlib
    h = hashlib.blake2s(prefix_words.tobytes() + nonce.to_bytes(8, "little")).digest()
    return int.from_bytes(h[:16], "little") % (1 << bits) == 0


def test_grind_stwo_family():
    # small bits: the answer lies in hi = 0, where both searches see 0, 1, 2, ...
    for bits in (4, 9, 14):
        ch = po.PyChannel()
        ch.mix_digest([bits, 9, 8, 7, 6, 5, 4, 3])
        assert ch.grind_stwo(bits) == ch.grind(bits)
    # 21 bits on this state: no low < 2^20 passes for hi = 0, so the answer
    # is (1 << 32) | low; checked with hashlib over every earlier nonce
    ch = po.PyChannel()
    ch.mix_digest([1, 1, 2, 3, 4, 5, 6, 7])
    n = ch.grind_stwo(21)
    pre = ch.pow_prefix(21)
    assert n >> 32 == 1 and _pow_ok(pre, 21, n)
    assert not any(_pow_ok(pre, 21, low) for low in range(1 << 20))
    assert not any(_pow_ok(pre, 21, (1 << 32) | low) for low in range(n & 0xFFFFFFFF))


def _tz_hashlib(prefix_words, nonce, m31_out):
    import hashlib
    h = hashlib.blake2s(np.asarray(prefix_words, np.uint32).tobytes() + nonce.to_bytes(8, "little")).digest()
    words = [int.from_bytes(h[4 * i:4 * i + 4], "little") for i in range(4)]
    if m31_out:
        words = [w % P for w in words]
    v = sum(w << (32 * i) for i, w in enumerate(words))
    return 128 if v == 0 else (v & -v).bit_length() - 1


def test_grind_m31_output_oracle():
    # Blake2sChannelGeneric<true> (grind_ops.rs:154-245): hash words reduced
    # mod P before the count; hashlib computes the same test independently
    rng = np.random.default_rng(64)
    differs = 0
    for _ in range(4):
        pre = rng.integers(0, P, 8, dtype=np.uint32)
        for n in list(range(200)) + [(1 << 32) | 5, (7 << 32) | 0xFFFFF]:
            a, b = po.pow_tz_words(pre, n, True), po.pow_tz_words(pre, n, False)
            assert a == _tz_hashlib(pre, n, True) and b == _tz_hashlib(pre, n, False)
            differs += a != b
    assert differs > 100  # the reduction changes the count whenever word 0 >= P
    # the search: valid, and every earlier nonce of hi = 0 fails
    for bits in (6, 11):
        pre = rng.integers(0, P, 8, dtype=np.uint32)
        n = po.grind_stwo_words(pre, bits, True)
        assert n < (1 << 20) and _tz_hashlib(pre, n, True) >= bits
        assert all(_tz_hashlib(pre, k, True) < bits for k in range(n))
    # the raw flavour over a channel's prefix is the channel-based search
    ch = po.PyChannel()
    ch.mix_digest([5, 4, 3, 2, 1, 0, 9, 8])
    assert po.grind_stwo_words(ch.pow_prefix(13), 13, False) == ch.grind_stwo(13)
    pm = ch.pow_prefix_m31(13)
    assert pm.dtype == np.uint32 and np.all(pm < P)


# ---- column ops: lifted accumulation (accumulate_lift.cu, accumulation_ops.rs)

def test_accumulate_lift_closed_form():
    rng = np.random.default_rng(61)
    curr = rng.integers(0, P, 8, dtype=np.uint32)
    col = rng.integers(0, P, 32, dtype=np.uint32)
    out = po.accumulate_lift(col, curr, 2)
    # i -> ((i >> 3) << 1) | (i & 1): rows 0..7 read curr[0], curr[1], curr[0], ...
    src = np.array([((i >> 3) << 1) | (i & 1) for i in range(32)])
    assert np.array_equal(out, (col.astype(np.int64) + curr[src]) % P)
    assert np.array_equal(po.accumulate_lift(col, col, 0), (2 * col.astype(np.int64)) % P)


def test_lift_and_accumulate_composes():
    # the loop's last column equals last + the sum of every smaller column
    # lifted directly (lifts by r1 then r2 equal the lift by r1 + r2)
    rng = np.random.default_rng(62)
    logs = [1, 3, 3, 5, 8]
    cols = [rng.integers(0, P, 1 << L, dtype=np.uint32) for L in logs]
    last = po.lift_and_accumulate(cols)[-1]
    L = logs[-1]
    want = cols[-1].astype(np.int64)
    for c, l in zip(cols[:-1], logs[:-1]):
        idx = np.array([((i >> (L - l + 1)) << 1) | (i & 1) for i in range(1 << L)])
        want = (want + c[idx]) % P
    assert np.array_equal(last, want)


# ---- half-coset natural <-> canonic bit-reversed order (src/circle.rs:192-220)

def _brev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def test_permute_hc_brt_literal_and_geometry():
    for L in range(0, 11):
        n = 1 << L
        src = np.random.default_rng(70 + L).integers(0, P, n, dtype=np.uint32)
        # the reference loop, restated here independently of the C oracle
        want = np.zeros(n, np.uint32)
        for k in range(n):
            cn = n - 1 - k // 2 if k & 1 else k // 2
            want[_brev(cn, L)] = src[k]
        got = po.permute_hc_to_canonic_brt(src)
        assert np.array_equal(got, want), L
        assert np.array_equal(po.permute_canonic_brt_to_hc_natural(got), src), L
        # both directions are one permutation (an involution)
        assert np.array_equal(po.permute_canonic_brt_to_hc_natural(src), got), L
        if L >= 1:
            # the geometry: point k of half_coset(L) is point cn(k) of the
            # canonic CircleDomain (half_odds(L - 1), then its conjugates)
            hc = po.coset(po.HALF_COSET, L)
            for k in range(n):
                cn = n - 1 - k // 2 if k & 1 else k // 2
                assert po.coset_at(hc, k) == po.stwo_domain_point(L, cn)


def test_permute_then_stwo_interpolate_matches_direct_evaluation():
    # values of a polynomial at the half-coset points, in natural order, go
    # through the permutation into the stwo NTT's input order: interpolate
    # then gives back the coefficients (the reference's
    # test_permute_then_stwo_intt_eval_domain, src/ntt.rs:451-500); from
    # 8 points up, where the stwo NTT is the CircleDomain transform
    for L in (3, 5, 8):
        n = 1 << L
        coeffs = np.random.default_rng(80 + L).integers(0, P, n, dtype=np.uint32)
        hc = po.coset(po.HALF_COSET, L)
        vals = np.zeros(n, np.uint32)
        for k in range(n):
            x, y = po.coset_at(hc, k)
            v = po.stwo_eval_at_point(coeffs, (x, 0, 0, 0), (y, 0, 0, 0))
            assert list(v[1:]) == [0, 0, 0]
            vals[k] = v[0]
        assert np.array_equal(po.stwo_interpolate(po.permute_hc_to_canonic_brt(vals)), coeffs), L


# ---- the OODS step (src/oods.rs): point from the channel, z_next, values, mix

def _qm(v):
    return tuple(int(x) for x in v)


def _on_circle(x, y):
    return pyref.qm_add(pyref.qm_mul(x, x), pyref.qm_mul(y, y)) == (1, 0, 0, 0)


def _channel(seed):
    ch = po.PyChannel()
    ch.mix_digest([seed, 3, 1, 4, 1, 5, 9, 2])
    return ch


def test_oods_sample_oracle():
    rng = np.random.default_rng(66)
    cols = [rng.integers(0, P, 1 << L, dtype=np.uint32) for L in (0, 1, 3, 6, 6, 9)]
    for trace_log in (0, 3, 9, 20, 31):
        ch = _channel(trace_log)
        pts, vals = ch.oods_sample(cols, trace_log)
        zx, zy, nx, ny = (_qm(pts[4 * i:4 * i + 4]) for i in range(4))
        assert _on_circle(zx, zy) and _on_circle(nx, ny)
        # z is the rational point of the drawn t: y = t (1 + x); the
        # reference channel here is pyref's hashlib restatement
        ref = pyref.Channel()
        ref.mix_digest([trace_log, 3, 1, 4, 1, 5, 9, 2])
        t = ref.draw_felt()
        assert zy == pyref.qm_mul(t, pyref.qm_add((1, 0, 0, 0), zx))
        # z_next: z moved by the 2^trace_log subgroup's generator, so
        # 2^trace_log steps come back to z (trace_log 0: z itself)
        if trace_log <= 9:
            gx, gy = po.coset_at(po.coset(po.SUBGROUP, trace_log), 1) if trace_log else (1, 0)
            x, y = zx, zy
            for _ in range(1 << trace_log):
                x, y = (pyref.qm_sub(pyref.qm_scale(x, gx), pyref.qm_scale(y, gy)),
                        pyref.qm_add(pyref.qm_scale(x, gy), pyref.qm_scale(y, gx)))
                if (x, y) == (zx, zy):
                    break
            assert (x, y) == (zx, zy)
            if trace_log:
                x1 = pyref.qm_sub(pyref.qm_scale(zx, gx), pyref.qm_scale(zy, gy))
                y1 = pyref.qm_add(pyref.qm_scale(zx, gy), pyref.qm_scale(zy, gx))
                assert (x1, y1) == (nx, ny)
            else:
                assert (nx, ny) == (zx, zy)
        # values: every column at z, then every column at z_next
        n = len(cols)
        for k, c in enumerate(cols):
            assert _qm(vals[4 * k:4 * k + 4]) == po.stwo_eval_at_point(c, zx, zy)
            assert _qm(vals[4 * (n + k):4 * (n + k) + 4]) == po.stwo_eval_at_point(c, nx, ny)
        # the channel: draw, then mix all the values in that order
        ref.mix_felts([_qm(vals[4 * i:4 * i + 4]) for i in range(2 * n)])
        assert ref.state == ch.state and ch.c.n_draws == 0
    # no columns: the draw only
    ch, ref = _channel(1), pyref.Channel()
    ref.mix_digest([1, 3, 1, 4, 1, 5, 9, 2])
    ch.oods_sample([], 5)
    ref.draw_felt()
    assert ch.state == ref.state and ch.c.n_draws == ref.n_draws
