"""Graph replay and the continuous-batching server (vx_server_*).

Every proof produced through a HIP graph or through the server, alone or in a
mix of concurrent requests, must equal byte for byte the proof of the same
request made alone by the kernel-by-kernel prover; that prover in turn equals
the oracle (test_gpu_prove.py)."""
import hashlib
import random

import pytest

import pyoracle as po
import vortexstark_amd as vx

pytestmark = pytest.mark.gpu


def alone(log_n, a, b, bb=2, pw=0):
    p = vx.Prover(log_n, bb, 80, pw, graphs=False)
    out = p.prove(a, b)
    p.close()
    return out


@pytest.mark.parametrize("log_n,bb,pw", [(8, 2, 0), (14, 2, 0), (12, 1, 12), (18, 2, 0)])
def test_graph_replays_equal_eager(log_n, bb, pw):
    g = vx.Prover(log_n, bb, 80, pw, graphs=True)
    e = vx.Prover(log_n, bb, 80, pw, graphs=False)
    for i, (a, b) in enumerate([(1, 1), (5, 7), (1, 1), (123456789, 42), (2, 3)]):
        print("replay", log_n, bb, pw, i, flush=True)
        want = e.prove(a, b)            # eager work between the replays
        got = g.prove(a, b)
        assert got == want, (log_n, i)
    assert g.prove(3, 4) == po.prove(3, 4, log_n, blowup_bits=bb, pow_bits=pw, omp=True)
    g.close()
    e.close()


def test_timed_proof_after_graph_replays():
    p = vx.Prover(12)
    a = p.prove(1, 2)
    p.enqueue(1, 2, timed=True)
    assert p.wait() == a
    assert p.stage_times()["total"] > 0
    assert p.prove(1, 2) == a
    p.close()


def mixed_requests(n, seed, sizes=(8, 10, 12, 14, 16), pw=(0, 10)):
    rng = random.Random(seed)
    reqs = []
    for _ in range(n):
        reqs.append(vx.request(rng.choice(sizes), a=rng.randrange(vx.P), b=rng.randrange(vx.P),
                               pow_bits=rng.choice(pw)))
    return reqs


@pytest.mark.parametrize("slots", [1, 4])
def test_server_mix_equals_alone(slots):
    reqs = mixed_requests(24, seed=slots)
    srv = vx.Server(n_slots=slots)
    tickets = [srv.submit(r) for r in reqs]
    got = [srv.wait(t) for t in tickets]
    srv.close()
    for r, (proof, (ts, t0, t1)) in zip(reqs, got):
        assert ts <= t0 <= t1
        assert proof == alone(r.log_n, r.a, r.b, r.blowup_bits, r.pow_bits), (r.log_n, r.a, r.b, r.pow_bits)
        assert vx.verify(proof)[0]


def test_server_run_trace_equals_alone():
    reqs = mixed_requests(40, seed=7, sizes=(10, 12, 14, 16, 18))
    rng = random.Random(1)
    t, arrivals = 0.0, []
    for _ in reqs:
        t += rng.expovariate(2000.0)
        arrivals.append(t)
    srv = vx.Server(n_slots=4)
    for k in {(r.log_n, r.pow_bits) for r in reqs}:
        srv.prepare(vx.request(k[0], pow_bits=k[1]))
    proofs, times = srv.run_trace(reqs, arrivals)
    srv.close()
    for r, p, (ts, t0, t1), at in zip(reqs, proofs, times, arrivals):
        assert at - 1e-3 <= ts <= t0 <= t1
        assert p == alone(r.log_n, r.a, r.b, r.blowup_bits, r.pow_bits)


def test_server_evicts_to_fit_budget():
    # budget for about one log_n=16 class: switching classes must evict
    budget = 3 * vx.Prover(16).device_bytes
    srv = vx.Server(n_slots=2, mem_budget=budget)
    reqs = [vx.request(n, a=n, b=n + 1) for n in (16, 14, 16, 12, 16, 14)]
    tickets = [srv.submit(r) for r in reqs]
    for r, t in zip(reqs, tickets):
        proof, _ = srv.wait(t)
        assert proof == alone(r.log_n, r.a, r.b)
    assert srv.resident_bytes <= budget
    srv.close()


def test_server_two_slots_tight_budget_alternating():
    # Regression test of the eviction / copy-out order: two slots and a budget
    # that holds one 2^16 class, classes alternating, all queued at once, so
    # proofs finish while a request of a new class waits for memory. A prover
    # whose finished proof is not yet copied out must not be evicted.
    one = vx.Prover(16)
    budget = int(2.2 * one.device_bytes)
    one.close()
    srv = vx.Server(n_slots=2, mem_budget=budget)
    classes = [(16, 0), (14, 0), (16, 6), (14, 6)]
    reqs = [vx.request(ln, a=i + 1, b=3 * i + 2, pow_bits=pw) for i, (ln, pw) in enumerate(classes * 4)]
    tickets = [srv.submit(r) for r in reqs]
    got = [srv.wait(t)[0] for t in tickets]
    st = srv.stats
    resident = srv.resident_bytes
    srv.close()
    print("server stats", st, "resident", resident, "budget", budget, flush=True)
    assert st["evicted"] > 0
    assert resident <= budget
    for r, proof in zip(reqs, got):
        assert proof == alone(r.log_n, r.a, r.b, r.blowup_bits, r.pow_bits), (r.log_n, r.a, r.pow_bits)


def test_server_rejects_what_cannot_fit():
    srv = vx.Server(n_slots=1, mem_budget=1 << 20)
    t = srv.submit(vx.request(16))
    with pytest.raises(vx.VortexError, match="budget"):
        srv.wait(t)
    # the server keeps serving what fits
    srv.close()
    with pytest.raises(vx.VortexError):
        srv2 = vx.Server(n_slots=1)
        srv2.submit(vx.request(16, a=vx.P))  # non-canonical input


def test_serving_mix_classes_match_oracle_digests(golden):
    # Every class of the serving workload (Fibonacci log_n 14..22 and the
    # 34-column AIR log_n 14..20, blowup 4, 80 queries, PoW 20), each request
    # twice, streamed in shuffled order on a Poisson trace through a 16-slot
    # server: every proof must equal the oracle's (SHA-256 committed by
    # tests/golden/gen_golden.py --serving).
    rows = golden["serving_e"]
    assert len(rows) == 16
    reqs, want = [], []
    for r in rows:
        if r["air"] == 0:
            q = vx.request(r["log_n"], a=r["a"], b=r["b"], blowup_bits=r["blowup_bits"], pow_bits=r["pow_bits"])
        else:
            q = vx.request(r["log_n"], air=vx.AIR_D2, seed=r["seed"], n_cols=r["n_cols"], n_cons=r["n_cons"],
                           air_seed=r["air_seed"], blowup_bits=r["blowup_bits"], pow_bits=r["pow_bits"])
        reqs.append(q)
        want.append(r["sha256"])
    order = list(range(len(reqs))) * 2
    rng = random.Random(5)
    rng.shuffle(order)
    t, arrivals = 0.0, []
    for _ in order:
        t += rng.expovariate(400.0)
        arrivals.append(t)
    srv = vx.Server(n_slots=16)
    proofs, _ = srv.run_trace([reqs[i] for i in order], arrivals)
    srv.close()
    for k, (i, proof) in enumerate(zip(order, proofs)):
        print("serving request", k, "class", rows[i]["air"], rows[i]["log_n"], flush=True)
        assert hashlib.sha256(proof).hexdigest() == want[i], (rows[i]["air"], rows[i]["log_n"])
