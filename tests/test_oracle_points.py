"""Single outputs of the twisted circle NTT (or_ntt_evaluate_at), the
oracle's O(n)-per-output evaluation used to check the log_n = 30 proof's
openings: against the full transform at every output for three coset kinds
and several zero-padding factors, against the LDE, and through the roundtrip
to the trace values."""
import numpy as np
import pytest

import pyoracle as po

P = po.P


@pytest.mark.parametrize("kind", [po.SUBGROUP, po.HALF_COSET, po.HALF_ODDS])
@pytest.mark.parametrize("log_n,extra", [(1, 0), (2, 0), (2, 2), (3, 1), (5, 0), (6, 3), (9, 2), (12, 1)])
def test_points_equal_full_transform(kind, log_n, extra):
    rng = np.random.default_rng(log_n * 10 + extra + 100 * kind)
    dom_log = log_n + extra
    c = rng.integers(0, P, size=1 << log_n, dtype=np.uint32)
    padded = np.zeros(1 << dom_log, np.uint32)
    padded[:c.size] = c
    dom = po.coset(kind, dom_log)
    full = po.evaluate(padded, dom)
    N = 1 << dom_log
    idx = np.arange(N) if N <= 1024 else np.sort(rng.choice(N, size=300, replace=False))
    got = po.evaluate_at(c, dom, idx)
    assert np.array_equal(got, full[idx])


def test_points_single_threaded_equal_openmp():
    rng = np.random.default_rng(3)
    c = rng.integers(0, P, size=1 << 18, dtype=np.uint32)
    dom = po.eval_domain(19)
    idx = rng.choice(1 << 19, size=40, replace=False)
    assert np.array_equal(po.evaluate_at(c, dom, idx, omp=False), po.evaluate_at(c, dom, idx, omp=True))


@pytest.mark.parametrize("log_n,bb", [(4, 1), (8, 2), (11, 1), (17, 2)])
def test_points_of_the_lde(log_n, bb):
    # the LDE of a trace = the eval-domain transform of its interpolant
    rng = np.random.default_rng(log_n)
    t = rng.integers(0, P, size=1 << log_n, dtype=np.uint32)
    want = po.lde(t, bb)
    coef = po.interpolate(t, po.coset(po.HALF_COSET, log_n))
    N = 1 << (log_n + bb)
    idx = np.concatenate([[0, 1, 2, 3, N - 1, N - 2], rng.choice(N, size=min(60, N), replace=False)])
    assert np.array_equal(po.evaluate_at(coef, po.eval_domain(log_n + bb), idx), want[idx])


def test_points_roundtrip_to_trace():
    log_n = 16
    t = po.fibonacci_trace(1, 12345, log_n)
    coef = po.interpolate(t, po.coset(po.HALF_COSET, log_n))
    idx = np.array([0, 1, 2, 5, 1000, (1 << log_n) - 1])
    assert np.array_equal(po.evaluate_at(coef, po.coset(po.HALF_COSET, log_n), idx), t[idx])


def test_points_catch_a_changed_coefficient():
    # every output depends on every coefficient (all factors are nonzero),
    # so one changed coefficient changes any output checked
    rng = np.random.default_rng(9)
    c = rng.integers(0, P, size=1 << 10, dtype=np.uint32)
    dom = po.eval_domain(12)
    idx = rng.choice(1 << 12, size=8, replace=False)
    base = po.evaluate_at(c, dom, idx)
    bad = c.copy()
    bad[777] = (int(bad[777]) + 1) % P
    assert np.all(po.evaluate_at(bad, dom, idx) != base)
