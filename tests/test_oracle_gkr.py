"""The oracle's GKR operations (cuda/gkr.cu restated in oracle/vx_oracle.c),
pinned by what the mathematics fixes rather than by retyping the kernels:

- folding the first variable is the affine map r -> (1 - r) lo + r hi, and
  folding every variable of an MLE evaluates it: sum_x f(x) eq(x, r);
- gen_eq_evals is v times the eq polynomial (closed form, entries summing
  to v), with y[0] on the index's top bit, the same convention folding uses;
- a grand-product layer keeps the product of all elements, a LogUp layer
  the sum of its fractions (checked with field inverses), down to one
  element;
- the round sums equal the eq-weighted sums of the next layer of the layer
  folded at t = 0 and t = 2 (the univariate round polynomial at those
  points), and a brute-force sum in pure Python.

The reference has no GKR test vectors; field arithmetic here is the
independent pure-Python restatement (tests/golden/pyref.py)."""
import random

import numpy as np
import pytest

import pyoracle as po
import pyref as pr

P = pr.P
ONE, ZERO = (1, 0, 0, 0), (0, 0, 0, 0)


def rq(rng):
    return tuple(rng.randrange(P) for _ in range(4))


def flat(qs):
    return np.array([w for q in qs for w in q], np.uint32)


def unflat(a):
    return [tuple(int(x) for x in a[4 * i:4 * i + 4]) for i in range(len(a) // 4)]


def qm_inv(x):
    # the multiplicative group of QM31 has order P^4 - 1
    e, r, b = P ** 4 - 2, ONE, x
    while e:
        if e & 1:
            r = pr.qm_mul(r, b)
        b = pr.qm_mul(b, b)
        e >>= 1
    return r


def eq_weight(x, k, r):
    """prod_j (bit (k-1-j) of x ? r[j] : 1 - r[j])."""
    w = ONE
    for j in range(k):
        w = pr.qm_mul(w, r[j] if (x >> (k - 1 - j)) & 1 else pr.qm_sub(ONE, r[j]))
    return w


@pytest.mark.parametrize("secure", [False, True])
def test_fix_first_variable_is_the_affine_fold(secure):
    rng = random.Random(1 + secure)
    n = 64
    vals = [rq(rng) if secure else (rng.randrange(P), 0, 0, 0) for _ in range(n)]
    col = flat(vals) if secure else np.array([v[0] for v in vals], np.uint32)
    lo, hi = vals[:n // 2], vals[n // 2:]
    assert unflat(po.gkr_fix_first_variable(col, ZERO, secure)) == lo
    assert unflat(po.gkr_fix_first_variable(col, ONE, secure)) == hi
    for _ in range(3):
        r = rq(rng)
        want = [pr.qm_add(pr.qm_mul(pr.qm_sub(ONE, r), a), pr.qm_mul(r, b)) for a, b in zip(lo, hi)]
        assert unflat(po.gkr_fix_first_variable(col, r, secure)) == want


@pytest.mark.parametrize("secure", [False, True])
def test_folding_every_variable_evaluates_the_mle(secure):
    rng = random.Random(7 + secure)
    k = 6
    vals = [rq(rng) if secure else (rng.randrange(P), 0, 0, 0) for _ in range(1 << k)]
    r = [rq(rng) for _ in range(k)]
    col = flat(vals) if secure else np.array([v[0] for v in vals], np.uint32)
    cur = po.gkr_fix_first_variable(col, r[0], secure)
    for j in range(1, k):
        cur = po.gkr_fix_first_variable(cur, r[j], True)
    got = unflat(cur)
    assert len(got) == 1
    brute = ZERO
    for x, f in enumerate(vals):
        brute = pr.qm_add(brute, pr.qm_mul(f, eq_weight(x, k, r)))
    assert got[0] == brute
    # the same value as the dot product with gen_eq_evals(r, 1)
    eq = unflat(po.gkr_gen_eq_evals(flat(r), ONE))
    dot = ZERO
    for f, e in zip(vals, eq):
        dot = pr.qm_add(dot, pr.qm_mul(f, e))
    assert dot == brute


@pytest.mark.parametrize("k", [0, 1, 2, 5, 7])
def test_gen_eq_evals_closed_form(k):
    rng = random.Random(100 + k)
    y, v = [rq(rng) for _ in range(k)], rq(rng)
    got = unflat(po.gkr_gen_eq_evals(flat(y) if k else np.zeros(0, np.uint32), v))
    assert len(got) == 1 << k
    for x in range(1 << k):
        assert got[x] == pr.qm_mul(v, eq_weight(x, k, y))
    total = ZERO
    for e in got:
        total = pr.qm_add(total, e)
    assert total == v  # sum_x eq(x, y) = 1


def test_gen_eq_evals_pass_definition():
    rng = random.Random(3)
    cur = 8
    buf = [rq(rng) for _ in range(cur)] + [ZERO] * cur
    y = rq(rng)
    got = unflat(po.gkr_gen_eq_evals_pass(flat(buf), y, cur))
    for j in range(cur):
        t = pr.qm_mul(buf[j], y)
        assert got[cur + j] == t and got[j] == pr.qm_sub(buf[j], t)


def test_grand_product_layers_keep_the_product():
    rng = random.Random(11)
    vals = [rq(rng) for _ in range(32)]
    prod = ONE
    for v in vals:
        prod = pr.qm_mul(prod, v)
    layer = flat(vals)
    n = 32
    first = True
    while n > 1:
        out, den = po.gkr_next_layer(po.GKR_GRAND_PRODUCT, layer, None, n)
        assert den is None
        got = unflat(out)
        if first:
            src = unflat(layer)
            assert got == [pr.qm_mul(src[2 * i], src[2 * i + 1]) for i in range(n // 2)]
            first = False
        layer, n = out, n // 2
    assert unflat(layer) == [prod]


def frac_sum(nums, dens):
    s = ZERO
    for a, d in zip(nums, dens):
        s = pr.qm_add(s, pr.qm_mul(a, qm_inv(d)))
    return s


@pytest.mark.parametrize("kind", [po.GKR_LOGUP_GENERIC, po.GKR_LOGUP_MULT, po.GKR_LOGUP_SINGLES])
def test_logup_layers_keep_the_fraction_sum(kind):
    rng = random.Random(20 + kind)
    n = 16
    dens = [rq(rng) for _ in range(n)]
    if kind == po.GKR_LOGUP_GENERIC:
        nums = [rq(rng) for _ in range(n)]
        num_arr = flat(nums)
    elif kind == po.GKR_LOGUP_MULT:
        m = [rng.randrange(P) for _ in range(n)]
        nums = [(x, 0, 0, 0) for x in m]
        num_arr = np.array(m, np.uint32)
    else:
        nums = [ONE] * n
        num_arr = None
    want = frac_sum(nums, dens)
    on, od = po.gkr_next_layer(kind, num_arr, flat(dens), n)
    gn, gd = unflat(on), unflat(od)
    for i in range(n // 2):  # the definition of a pair's sum
        assert gd[i] == pr.qm_mul(dens[2 * i], dens[2 * i + 1])
        assert gn[i] == pr.qm_add(pr.qm_mul(nums[2 * i], dens[2 * i + 1]), pr.qm_mul(nums[2 * i + 1], dens[2 * i]))
    # the following layers are generic; the fraction sum never changes
    n //= 2
    while n > 1:
        on, od = po.gkr_next_layer(po.GKR_LOGUP_GENERIC, on, od, n)
        n //= 2
    (nn,), (dd,) = unflat(on), unflat(od)
    assert pr.qm_mul(nn, qm_inv(dd)) == want


def layer_inputs(kind, rng, length):
    dens = [rq(rng) for _ in range(length)]
    if kind == po.GKR_GRAND_PRODUCT:
        nums = [rq(rng) for _ in range(length)]
        return nums, flat(nums), None, None
    if kind == po.GKR_LOGUP_GENERIC:
        nums = [rq(rng) for _ in range(length)]
        return nums, flat(nums), dens, flat(dens)
    if kind == po.GKR_LOGUP_MULT:
        m = [rng.randrange(P) for _ in range(length)]
        return [(x, 0, 0, 0) for x in m], np.array(m, np.uint32), dens, flat(dens)
    return [ONE] * length, None, dens, flat(dens)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("n_terms", [1, 2, 8])
def test_sum_poly_is_the_round_polynomial_at_0_and_2(kind, n_terms):
    rng = random.Random(1000 * kind + n_terms)
    length = 4 * n_terms
    nums, num_arr, dens, den_arr = layer_inputs(kind, rng, length)
    eq = [rq(rng) for _ in range(n_terms)]
    lam = rq(rng) if kind else ZERO
    s0, s2 = po.gkr_sum_poly(kind, flat(eq), num_arr, den_arr, lam if kind else None, n_terms)
    # brute force: the layer's elements at t are (1 - t) lo + t hi
    for t, got in ((0, s0), (2, s2)):
        tq = (t, 0, 0, 0)

        def at(vals, j):
            return pr.qm_add(pr.qm_mul(pr.qm_sub(ONE, tq), vals[j]), pr.qm_mul(tq, vals[j + 2 * n_terms]))
        acc = ZERO
        for i in range(n_terms):
            a, b = at(nums, 2 * i), at(nums, 2 * i + 1)
            if kind == po.GKR_GRAND_PRODUCT:
                f = pr.qm_mul(a, b)
            else:
                da, db = at(dens, 2 * i), at(dens, 2 * i + 1)
                f = pr.qm_add(pr.qm_add(pr.qm_mul(a, db), pr.qm_mul(b, da)), pr.qm_mul(lam, pr.qm_mul(da, db)))
            acc = pr.qm_add(acc, pr.qm_mul(eq[i], f))
        assert tuple(int(x) for x in got) == acc
    # composition of the oracle's other operations: fold the first variable
    # at t, take the next layer, weight by eq
    for t, got in ((0, s0), (2, s2)):
        tq = (t, 0, 0, 0)
        if kind == po.GKR_GRAND_PRODUCT:
            fn = po.gkr_fix_first_variable(num_arr, tq, True)
            out, _ = po.gkr_next_layer(po.GKR_GRAND_PRODUCT, fn, None, 2 * n_terms)
            vals = unflat(out)
        else:
            fd = po.gkr_fix_first_variable(den_arr, tq, True)
            if kind == po.GKR_LOGUP_SINGLES:
                on, od = po.gkr_next_layer(po.GKR_LOGUP_SINGLES, None, fd, 2 * n_terms)
            else:
                fn = po.gkr_fix_first_variable(num_arr, tq, kind == po.GKR_LOGUP_GENERIC)
                on, od = po.gkr_next_layer(po.GKR_LOGUP_GENERIC, fn, fd, 2 * n_terms)
            vals = [pr.qm_add(a, pr.qm_mul(lam, d)) for a, d in zip(unflat(on), unflat(od))]
        acc = ZERO
        for e, v in zip(eq, vals):
            acc = pr.qm_add(acc, pr.qm_mul(e, v))
        assert tuple(int(x) for x in got) == acc
