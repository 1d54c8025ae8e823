"""The circle NTT's tuned plans: every plan (built-in, tiles in launch
order, strided bits over one more pass, both) gives the oracle's transform,
in place and out of place, in both bases, forced per shape through a tuner
cache file; the tuner measures a shape on first use and records its choice."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from gpu_util import dev, empty, host, ptr, sync
from test_gpu_kernels import _itables, _tables, hp, rnd
from vortexstark_amd._lib import call

pytestmark = pytest.mark.gpu
NULL = None


def force(tmp_path, lines):
    f = tmp_path / "tune.txt"
    f.write_text("".join(l + "\n" for l in lines))
    call("vx_tuner_configure", 1, str(f).encode())


def ntt_entries():
    n = C.c_uint32()
    call("vx_tuner_ntt_entries", None, 0, C.byref(n))
    out = (C.c_uint32 * (8 * max(n.value, 1)))()
    call("vx_tuner_ntt_entries", out, n.value, C.byref(n))
    return [tuple(out[8 * i:8 * i + 8]) for i in range(n.value)]


PLANS = [(0, 0), (0, 1), (1, 0), (1, 1)]


@pytest.mark.parametrize("log_n", [16, 20, 22])
@pytest.mark.parametrize("split,order", PLANS)
def test_every_plan_matches_the_oracle(tmp_path, log_n, split, order):
    print("ntt plan", split, order, "log_n", log_n, flush=True)
    force(tmp_path, [f"ntt 1 {log_n} {log_n} 1 0 {split} {order}", f"ntt 0 {log_n} {log_n} 1 0 {split} {order}"])
    try:
        c = po.coset(po.HALF_COSET, log_n)
        v = rnd(1 << log_n, 300 + log_n)
        line, circ, offs, sizes = _tables(c)
        d = dev(v)
        call("vx_circle_ntt_evaluate", ptr(d), ptr(line), ptr(circ), hp(offs), hp(sizes), log_n - 1, 1 << log_n, NULL)
        assert np.array_equal(host(d), po.evaluate(v, c))
        iline, icirc, ioffs, isizes = _itables(c)
        call("vx_circle_ntt_interpolate", ptr(d), ptr(iline), ptr(icirc), hp(ioffs), hp(isizes), log_n - 1,
             1 << log_n, NULL)
        assert np.array_equal(host(d), v)
    finally:
        force(tmp_path, [f"ntt 1 {log_n} {log_n} 1 0 0 0", f"ntt 0 {log_n} {log_n} 1 0 0 0"])


@pytest.mark.parametrize("split,order", PLANS)
def test_every_plan_lde_matches_the_oracle(tmp_path, split, order):
    log_n, blowup = 20, 2
    le = log_n + blowup
    force(tmp_path, [f"ntt 1 {le} {log_n} 1 0 {split} {order}"])
    try:
        c = po.eval_domain(le)
        coef = rnd(1 << log_n, 77)
        line, circ, offs, _ = _tables(c)
        dc, out = dev(coef), empty(1 << le)
        call("vx_circle_ntt_evaluate_lde", ptr(out), ptr(dc), 1 << log_n, ptr(line), ptr(circ), hp(offs), le - 1,
             1 << le, NULL)
        padded = np.zeros(1 << le, np.uint32)
        padded[:1 << log_n] = coef
        assert np.array_equal(host(out), po.evaluate(padded, c))
    finally:
        force(tmp_path, [f"ntt 1 {le} {log_n} 1 0 0 0"])


@pytest.mark.parametrize("split,order", PLANS)
def test_every_plan_stwo_basis(tmp_path, split, order):
    log_n = 19
    n = 1 << log_n
    force(tmp_path, [f"ntt 1 {log_n} {log_n} 1 1 {split} {order}", f"ntt 0 {log_n} {log_n} 1 1 {split} {order}"])
    try:
        fwd, inv = po.stwo_twiddles(log_n)
        tw, itw = dev(fwd), dev(inv)
        x = rnd(n, 5)
        d = dev(x)
        call("vx_stwo_ntt_evaluate", ptr(d), ptr(tw), n, NULL)
        assert np.array_equal(host(d), po.stwo_evaluate(x, fwd, omp=True))
        call("vx_stwo_ntt_interpolate", ptr(d), ptr(itw), n, NULL)
        assert np.array_equal(host(d), x)
    finally:
        force(tmp_path, [f"ntt 1 {log_n} {log_n} 1 1 0 0", f"ntt 0 {log_n} {log_n} 1 1 0 0"])


def test_tuner_measures_a_new_shape():
    # a shape no other test uses: measured on first use, result unchanged
    log_n, blowup = 21, 1
    le = log_n + blowup
    c = po.eval_domain(le)
    coef = rnd(1 << log_n, 91)
    line, circ, offs, _ = _tables(c)
    dc, out = dev(coef), empty(1 << le)
    call("vx_circle_ntt_evaluate_lde", ptr(out), ptr(dc), 1 << log_n, ptr(line), ptr(circ), hp(offs), le - 1, 1 << le,
         NULL)
    sync()
    padded = np.zeros(1 << le, np.uint32)
    padded[:1 << log_n] = coef
    assert np.array_equal(host(out), po.evaluate(padded, c))
    keys = [e[1:6] for e in ntt_entries()]
    assert (1, le, log_n, 1, 0) in keys
