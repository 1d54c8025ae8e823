"""The synthetic multi-column AIR (KRAK v3) on the CPU: the C oracle against
the independent Python restatement (tests/golden/pyref.py) and its own
committed digests, the definition's closed forms, the oracle verifier and the
product's host verifier (vx_verify, host code in libvortex_hip.so) on
accepted and tampered proofs."""
import hashlib
import struct

import numpy as np
import pytest

import pyoracle as po
import pyref
import vortexstark_amd as vx

P = po.P


def test_d2_trace_values_closed_form():
    # t_c[i] = mix64(seed + G ((c << 32 | i) + 1)) mod P
    for seed, c, i in [(0, 0, 0), (42, 3, 17), (2**64 - 1, 33, 2**20 - 1), (0xDEADBEEF, 16, 5)]:
        want = pyref.mix64((seed + pyref.GOLDEN * (((c << 32) | i) + 1)) & pyref.M64) % P
        assert po.lib().or_d2_value(seed, c, i) == want
    t = po.d2_trace(9, 5, 4)
    assert t.shape == (5, 16)
    assert all(int(t[c, i]) == pyref.d2_value(9, c, i) for c in range(5) for i in range(16))
    assert (t < P).all()


def test_d2_coefficients_are_splitmix_stream():
    cd = po.d2_coeffs(0xD2A1, 35)
    assert list(map(int, cd)) == pyref.d2_coeffs(0xD2A1, 35)
    # SplitMix64 reference value: the first output of seed 0 is 0xE220A8397B1DCDAF
    assert pyref.mix64(pyref.GOLDEN) == 0xE220A8397B1DCDAF
    assert int(po.d2_coeffs(0, 1)[0]) == 0xE220A8397B1DCDAF % P


def test_d2_quotient_brute_force():
    rng = np.random.default_rng(5)
    W, K, N = 7, 9, 16
    ev = rng.integers(0, P, size=(W, N), dtype=np.uint32)
    alpha = [3, 1, 4, 1]
    q = po.d2_quotient(ev, K, 11, alpha)
    cd = pyref.d2_coeffs(11, K)
    for i in range(N):
        want = pyref.d2_q([int(x) for x in ev[:, i]], [int(x) for x in ev[:, (i + 1) % N]], W, K, cd, tuple(alpha))
        assert tuple(int(q[c][i]) for c in range(4)) == want


def test_d2_single_constraint_reduces_to_scaled_product():
    # K = 1: Q = alpha (c t_0[i] t_1[i+1] + d t_7 mod W [i]) -- a dropped term or
    # a wrong row offset changes this value
    W, N = 3, 8
    ev = np.arange(1, W * N + 1, dtype=np.uint32).reshape(W, N)
    alpha = [2, 0, 0, 0]
    q = po.d2_quotient(ev, 1, 77, alpha)
    c, d = pyref.d2_coeffs(77, 1)
    for i in range(N):
        m = (c * int(ev[0, i]) * int(ev[1, (i + 1) % N]) + d * int(ev[7 % W, i])) % P
        assert int(q[0][i]) == 2 * m % P and int(q[1][i]) == 0


def test_d2_oracle_matches_python_restatement(golden):
    for g in golden["proofs_d2_py"]:
        data = po.d2_prove(g["seed"], g["log_n"], g["n_cols"], g["n_cons"], g["air_seed"], g["blowup_bits"],
                           g["n_queries"], g["pow_bits"])
        assert len(data) == g["len"] == po.d2_proof_size(g["log_n"], g["n_cols"], g["blowup_bits"], g["n_queries"])
        assert hashlib.sha256(data).hexdigest() == g["sha256"], g


def test_d2_oracle_digests_stable(golden):
    for g in golden["proofs_d2_c"]:
        if g["log_n"] > 14:
            continue
        data = po.d2_prove(g["seed"], g["log_n"], g["n_cols"], g["n_cons"], g["air_seed"], g["blowup_bits"],
                           g["n_queries"], g["pow_bits"], omp=True)
        assert hashlib.sha256(data).hexdigest() == g["sha256"], g


def test_serving_fixture_is_the_generated_request_list(golden):
    # the committed serving-mix digests belong to exactly the requests the
    # generator draws, and the two smallest classes reproduce from the oracle
    import gen_golden

    rows = golden["serving_e"]
    reqs = gen_golden.serving_requests_e()
    assert [{k: v for k, v in r.items() if k not in ("len", "sha256")} for r in rows] == reqs
    fib14 = next(r for r in rows if r["air"] == 0 and r["log_n"] == 14)
    data = po.prove(fib14["a"], fib14["b"], 14, blowup_bits=2, n_queries=80, pow_bits=20, omp=True)
    assert hashlib.sha256(data).hexdigest() == fib14["sha256"]
    d14 = next(r for r in rows if r["air"] == 1 and r["log_n"] == 14)
    data = po.d2_prove(d14["seed"], 14, 34, 35, 0xD2A1, 2, 80, 20, omp=True)
    assert hashlib.sha256(data).hexdigest() == d14["sha256"]
    ok, why = vx.verify(data, 2)
    assert ok, why


def words(b):
    return list(struct.unpack(f"<{len(b) // 4}I", b))


def pack(w):
    return struct.pack(f"<{len(w)}I", *w)


def d2_tampers(proof, W=34, nq=80):
    """Field-level tampers of a KRAK v3 proof (log_n from the header)."""
    w = words(proof)
    log_n, bb = w[2], w[5]
    le = log_n + bb
    G = w[13]
    troot0 = 14
    qroot = troot0 + 8 * G
    n_fri = w[qroot + 8]
    last = qroot + 9 + 8 * n_fri
    nqw = last + 1 + 32
    tdec = nqw + 1 + nq
    width0 = min(16, W)
    cases = {}

    def flip(name, i, delta=1):
        x = list(w)
        x[i] = (x[i] + delta) % P if x[i] < P else x[i] ^ 1
        cases[name] = pack(x)

    flip("trace_root", troot0 + 3)
    flip("quotient_root", qroot + 1)
    flip("fri_root", qroot + 9)
    flip("last_layer", last + 5)
    flip("query_index", nqw + 2)
    flip("trace_value_at_q", tdec + 1 + 3)
    flip("trace_value_next_row", tdec + 1 + nq * width0 + 2)
    x = list(w)
    x[11] ^= 1  # air_seed: the constraint check must fail
    cases["air_seed"] = pack(x)
    x = list(w)
    x[4] += 1   # constraint count
    cases["n_cons"] = pack(x)
    cases["truncated"] = proof[:-4]
    cases["trailing"] = proof + b"\0\0\0\0"
    return cases


@pytest.mark.parametrize("log_n,pw", [(4, 0), (6, 6)])
def test_d2_verifiers_accept_and_reject(log_n, pw):
    proof = po.d2_prove(3, log_n, 34, 35, 0xD2A1, 2, 80, pw)
    assert po.d2_verify(proof) == (True, "")
    assert vx.verify(proof) == (True, "")
    for name, bad in d2_tampers(proof).items():
        assert not po.d2_verify(bad)[0], name
        ok, why = vx.verify(bad)
        assert not ok, name
    ok, why = vx.verify(d2_tampers(proof)["air_seed"])
    assert "Constraint quotient mismatch" in why
    assert not vx.verify(proof, blowup_bits=1)[0]


def test_d2_proof_size_matches_layout():
    for log_n, W, bb in [(4, 34, 2), (6, 3, 1), (8, 17, 2), (10, 16, 3)]:
        assert vx.d2_proof_size(log_n, W, bb, 80) == po.d2_proof_size(log_n, W, bb, 80)
    assert len(po.d2_prove(1, 6, 17, 4, 5, 2, 80, 0)) == vx.d2_proof_size(6, 17, 2, 80)
