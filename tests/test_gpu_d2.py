"""GPU proofs of the synthetic multi-column AIR (KRAK v3) through the C-ABI:
byte-equal to the C oracle at small sizes and to the committed oracle
digests at full serving sizes; graph replays; the batching server on a mix
of multi-column and Fibonacci requests."""
import hashlib

import pytest

import pyoracle as po
import vortexstark_amd as vx

pytestmark = pytest.mark.gpu


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("log_n,W,K,air,bb,pw", [
    (4, 34, 35, 0xD2A1, 2, 0), (6, 34, 35, 0xD2A1, 2, 0), (8, 3, 2, 99, 1, 0), (8, 17, 4, 5, 2, 6),
    (10, 16, 40, 1, 2, 0), (12, 34, 35, 0xD2A1, 2, 10), (9, 1, 3, 7, 3, 0), (13, 48, 60, 3, 1, 0)])
def test_d2_proofs_match_oracle(log_n, W, K, air, bb, pw):
    print("d2", log_n, W, K, bb, pw, flush=True)
    p = vx.D2Prover(log_n, W, K, air, bb, 80, pw)
    for seed in (3, 0xFFFFFFFF12345678):
        got = p.prove(seed)
        want = po.d2_prove(seed, log_n, W, K, air, bb, 80, pw, omp=True)
        assert len(got) == len(want) == p.size
        assert got == want, (log_n, W, K, seed)
    assert vx.verify(got, bb) == (True, "")
    p.close()


@pytest.mark.parametrize("log_n", [10, 14, 16, 20])
def test_d2_full_size_digests(golden, log_n):
    g = next(x for x in golden["proofs_d2_c"] if x["log_n"] == log_n)
    p = vx.D2Prover(log_n, g["n_cols"], g["n_cons"], g["air_seed"], g["blowup_bits"], g["n_queries"], g["pow_bits"])
    got = p.prove(g["seed"])
    assert len(got) == g["len"]
    assert sha(got) == g["sha256"]
    assert vx.verify(got) == (True, "")
    p.close()


def test_d2_graph_replay_equals_eager():
    g = vx.D2Prover(12, pow_bits=8, graphs=True)
    e = vx.D2Prover(12, pow_bits=8, graphs=False)
    for seed in (1, 2, 1, 99, 2**63 + 5):
        assert g.prove(seed) == e.prove(seed)
    g.close()
    e.close()


def test_server_mixed_airs_equal_alone():
    reqs = []
    for i, ln in enumerate([10, 12, 8, 14, 10, 12, 9, 11]):
        reqs.append(vx.request(ln, air=vx.AIR_D2, seed=1000 + i, pow_bits=6))
        reqs.append(vx.request(ln + 1, a=i + 1, b=2 * i + 3, pow_bits=6))
    srv = vx.Server(n_slots=4)
    tickets = [srv.submit(r) for r in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    srv.close()
    for r, proof in zip(reqs, got):
        if r.air == vx.AIR_D2:
            want = po.d2_prove(r.seed, r.log_n, r.n_cols, r.n_cons, r.air_seed, r.blowup_bits, 80, r.pow_bits,
                               omp=True)
        else:
            want = po.prove(r.a, r.b, r.log_n, blowup_bits=r.blowup_bits, pow_bits=r.pow_bits, omp=True)
        assert proof == want, (r.air, r.log_n)
