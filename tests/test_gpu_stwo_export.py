"""stwo wire-format export of proofs made on the GPU (Fibonacci v1/v2, the
multi-column AIR v3, the batched columns v4): the proof bytes equal the
oracle's, the exported JSON is the export of the oracle's proof, and every
tree and FRI layer root rebuilds from its hash witness."""
import pytest

import pyoracle as po
import vortexstark_amd as vx
from vortexstark_amd import stwo_export as se

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_n,pow_bits", [(12, 0), (16, 8)])
def test_fibonacci_proof_export(log_n, pow_bits):
    print("fibonacci export log_n", log_n, "pow", pow_bits, flush=True)
    g = vx.prove(3, 5, log_n, 2, 80, pow_bits)
    assert g == po.prove(3, 5, log_n, 2, 80, pow_bits, omp=True)
    two = se.proof_to_stwo(g)
    se.verify_all_merkle_witnesses(two)
    assert two["config"]["pow_bits"] == pow_bits and len(two["query_indices"]) == 80


def test_d2_proof_export():
    print("d2 export log_n 10", flush=True)
    p = vx.D2Prover(10, 34, 35, 0xD2A1, 2, 80, 4)
    try:
        g = p.prove(9)
    finally:
        p.close()
    assert g == po.d2_prove(9, 10, 34, 35, 0xD2A1, 2, 80, 4, omp=True)
    two = se.proof_to_stwo(g)
    se.verify_all_merkle_witnesses(two)
    assert [m["n_cols"] for m in two["tree_meta"]] == [16, 16, 2, 4]


def test_batch_proof_export():
    logs = [14, 14, 15, 16]
    print("batch export", logs, flush=True)
    p = vx.BatchProver(logs, 2, 40, 4)
    try:
        g = p.prove(21)
    finally:
        p.close()
    assert g == po.batch_prove(21, logs, 2, 40, 4)
    two = se.proof_to_stwo(g)
    se.verify_all_merkle_witnesses(two)
    assert two["tree_meta"][0] == {"name": "lifted_trace", "n_cols": 4, "log_eval_size": 18}
