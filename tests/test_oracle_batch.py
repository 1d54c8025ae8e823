"""The batched-columns proof (KRAK v4) on the CPU: the C oracle against the
independent Python restatement (tests/golden/pyref.py) and its own digests;
the oracle verifier and the product's host verifier (vx_verify, host code)
on accepted and tampered proofs."""
import hashlib
import struct

import pytest

import pyoracle as po
import vortexstark_amd as vx


def test_batch_oracle_matches_python_restatement(golden):
    for g in golden["proofs_batch_py"]:
        data = po.batch_prove(g["seed"], g["log_sizes"], g["blowup_bits"], g["n_queries"], g["pow_bits"])
        assert len(data) == g["len"] == po.batch_proof_size(g["log_sizes"], g["blowup_bits"], g["n_queries"])
        assert hashlib.sha256(data).hexdigest() == g["sha256"], g


def test_batch_oracle_digest_stable(golden):
    g = next(x for x in golden["proofs_batch_c"] if x["n_cols"] == 12)
    data = po.batch_prove(g["seed"], g["log_sizes"], g["blowup_bits"], g["n_queries"], g["pow_bits"])
    assert hashlib.sha256(data).hexdigest() == g["sha256"]


def tampers(proof, nq):
    w = list(struct.unpack(f"<{len(proof) // 4}I", proof))
    C = w[2]
    troot = 9 + C
    qroot = troot + 8
    n_fri = w[qroot + 8]
    last = qroot + 9 + 8 * n_fri
    nqw = last + 1 + 32
    tdec = nqw + 1 + nq
    cases = {}

    def flip(name, i):
        x = list(w)
        x[i] = (x[i] + 1) % po.P
        cases[name] = struct.pack(f"<{len(x)}I", *x)

    flip("trace_root", troot + 2)
    flip("quotient_root", qroot)
    flip("fri_root", qroot + 9)
    flip("last_layer", last + 3)
    flip("query", nqw + 1)
    flip("opened_value", tdec + 1 + C // 2)
    flip("sibling_value", tdec + 1 + nq * C + 1)
    flip("log_size", 9)
    cases["truncated"] = proof[:-4]
    cases["trailing"] = proof + b"\0" * 4
    return cases


@pytest.mark.parametrize("logs,pw", [([2, 3, 3, 4], 0), ([3, 3, 5, 5, 5], 6)])
def test_batch_verifiers_accept_and_reject(logs, pw):
    proof = po.batch_prove(11, logs, 2, 16, pw)
    assert po.batch_verify(proof, 2, 16) == (True, "")
    assert vx.verify(proof, 2, 16) == (True, "")
    for name, bad in tampers(proof, 16).items():
        assert not po.batch_verify(bad, 2, 16)[0], name
        assert not vx.verify(bad, 2, 16)[0], name
    assert not vx.verify(proof, 1, 16)[0]


def test_batch_proof_size_matches_layout():
    for logs, bb in [([14, 15, 16], 2), ([14] * 20 + [17], 1), ([20, 22], 2)]:
        assert vx.batch_proof_size(logs, bb, 80) == po.batch_proof_size(logs, bb, 80)
