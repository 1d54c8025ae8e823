"""End-to-end parity of GPU proofs (vx_prover_* through the C-ABI).

Small sizes: proof bytes equal the oracle's bytes and the independent Python
restatement's digests. Full size (log_n = 20, 24, 25 at blowup 1, and 24 with
26 PoW bits): the SHA-256 of the proof equals the oracle digest committed in
tests/golden/golden.json (computed by oracle/liboracle_omp.so), and the
verifier accepts. log_n = 30 (blowup 1) is checked by properties only."""
import hashlib

import numpy as np
import pytest

import pyoracle as po
import vortexstark_amd as vx
from gpu_util import d2h
from krak import Krak
from test_oracle import tamper_cases

pytestmark = pytest.mark.gpu


def sha(b):
    return hashlib.sha256(b).hexdigest()


def test_proofs_match_independent_python_digests(golden):
    for g in golden["proofs_py"]:
        print("proof", g["log_n"], g["blowup_bits"], g["pow_bits"], flush=True)
        p = vx.Prover(g["log_n"], g["blowup_bits"], 80, g["pow_bits"])
        got = p.prove(g["a"], g["b"])
        assert len(got) == g["len"]
        assert sha(got) == g["sha256"], g
        p.close()


@pytest.mark.parametrize("log_n,bb,pw", [(10, 2, 0), (12, 2, 0), (14, 1, 0), (16, 2, 0), (16, 2, 10)])
def test_proofs_match_oracle(golden, log_n, bb, pw):
    g = next(x for x in golden["proofs_c"] if (x["log_n"], x["blowup_bits"], x["pow_bits"]) == (log_n, bb, pw))
    p = vx.Prover(log_n, bb, 80, pw)
    got = p.prove(g["a"], g["b"])
    if sha(got) != g["sha256"]:
        want = po.prove(g["a"], g["b"], log_n, blowup_bits=bb, pow_bits=pw, omp=True)
        kg, kw = Krak(got), Krak(want)
        first = next(i for i, (x, y) in enumerate(zip(kg.words, kw.words)) if x != y)
        pytest.fail(f"proof differs from the oracle at word {first} "
                    f"(trace root ok: {kg.root(kg.off_trace_root) == kw.root(kw.off_trace_root)})")
    assert vx.verify(got, bb) == (True, "")


@pytest.mark.parametrize("log_n,bb,pw", [(20, 2, 0), (24, 2, 0), (25, 1, 0), (24, 2, 26)])
def test_full_size_proof_matches_oracle_digest(golden, log_n, bb, pw):
    g = next((x for x in golden["proofs_c"] if (x["log_n"], x["blowup_bits"], x["pow_bits"]) == (log_n, bb, pw)),
             None)
    if g is None:
        pytest.skip("digest not in golden.json (generate with gen_golden.py --big)")
    p = vx.Prover(log_n, bb, 80, pw)
    got = p.prove(g["a"], g["b"])
    assert len(got) == g["len"]
    assert sha(got) == g["sha256"]
    assert vx.verify(got, bb) == (True, "")
    p.close()


def test_log30_blowup1_proof_properties(golden):
    """log_n = 30 at blowup 1: a 2^31-point evaluation domain, the largest the
    reference's u32 indexing admits. Too large for the oracle, so the checks
    are size-independent: the verifier accepts, the proof has the layout's
    length and KRAK v1 header, proving is deterministic, and another statement
    gives another proof."""
    ws = golden["workload_seed"]
    p = vx.Prover(30, 1)
    print("device bytes", p.device_bytes, flush=True)
    got = p.prove(ws["a0"], ws["a1"])
    assert len(got) == vx.proof_size(30, 1)
    k = Krak(got)
    assert (k.magic, k.version, k.log_n) == (0x4B52414B, 1, 30)
    assert vx.verify(got, 1) == (True, "")
    assert p.prove(ws["a0"], ws["a1"]) == got
    other = p.prove(ws["a0"], ws["a1"] ^ 1)
    assert other != got and vx.verify(other, 1) == (True, "")
    bad = bytearray(got)
    bad[len(bad) // 2] ^= 1
    assert not vx.verify(bytes(bad), 1)[0]
    p.close()


def test_intermediate_lde_matches_oracle():
    log_n, a, b = 12, 3, 5
    p = vx.Prover(log_n)
    p.prove(a, b)
    ptr, words = p.buffer(1)
    got = d2h(ptr, words)
    c = po.coset(po.HALF_COSET, log_n)
    coef = po.interpolate(po.fibonacci_trace(a, b, log_n), c)
    padded = np.zeros(words, np.uint32)
    padded[:1 << log_n] = coef
    assert np.array_equal(got, po.evaluate(padded, po.eval_domain(log_n + 2)))
    ptr0, w0 = p.buffer(0)
    assert np.array_equal(d2h(ptr0, w0), coef)
    p.close()


def test_prover_reuse_is_deterministic_and_input_sensitive():
    p = vx.Prover(14)
    a1 = p.prove(1, 1)
    b1 = p.prove(2, 3)
    a2 = p.prove(1, 1)
    assert a1 == a2 and a1 != b1
    assert a1 == po.prove(1, 1, 14, omp=True)
    assert b1 == po.prove(2, 3, 14, omp=True)
    p.close()


def test_gpu_proof_tamper_rejected():
    proof = vx.Prover(8).prove(1, 1)
    assert vx.verify(proof)[0]
    for name, bad in tamper_cases(proof).items():
        assert not vx.verify(bad)[0], name
        assert not po.verify(bad)[0], name


@pytest.mark.parametrize("log_n,pw", [(6, 8), (10, 14), (12, 18)])
def test_pow_v2_matches_oracle(log_n, pw):
    got = vx.Prover(log_n, 2, 80, pw).prove(7, 9)
    assert got == po.prove(7, 9, log_n, pow_bits=pw, omp=True)
    assert vx.verify(got)[0]


def test_stage_times_reported():
    p = vx.Prover(16)
    p.enqueue(1, 1, timed=True)
    proof = p.wait()
    t = p.stage_times()
    assert t["total"] > 0 and all(v >= 0 for v in t.values())
    assert vx.verify(proof)[0]
    p.close()
