"""stwo wire-format export (src/stwo_export.rs) of KRAK proofs, host code
in libvortex_hip.so (vx_export_stwo_json and the witness helpers).

Pinned by: the reference's own unit vectors for auth_paths_to_hash_witness
(src/stwo_export.rs:884-942) and dedup_fri_witness (:1592-1637); every
exported tree root, FRI first layer and inner layer rebuilt from its hash
witness by an independent hashlib restatement equals the commitment the
oracle prover wrote (a wrong witness order or a missing hash fails it);
the fri_witness values equal the opened values at unpaired siblings and
their count the reference's unpaired-position count (:1644-1693); the last
layer coefficients evaluated as a stwo LinePoly reproduce the last-layer
evaluations. (KRAK quotients take their neighbours in evaluation order,
cuda/constraint.cu:8-24, so the last layer carries no degree bound
and all of its coefficients may be nonzero.)"""
import hashlib
import json
import random

import numpy as np
import pytest

import pyoracle as po
import pyref
from vortexstark_amd import stwo_export as se
from vortexstark_amd._lib import VortexError

P = po.P


def h(n):
    return [n, 0, 0, 0, 0, 0, 0, 0]


# ---------------------------------------------------------------- unit vectors
def test_hash_witness_no_queries():
    assert se.auth_paths_to_hash_witness([], [], 5) == []


def test_hash_witness_single_query():
    assert se.auth_paths_to_hash_witness([0], [[h(1), h(2)]], 2) == [tuple(h(1)), tuple(h(2))]


def test_hash_witness_sibling_pair_deduplication():
    w = se.auth_paths_to_hash_witness([0, 1], [[h(1), h(2)], [h(1), h(2)]], 2)
    assert w == [tuple(h(2))]


def test_hash_witness_two_isolated_queries():
    w = se.auth_paths_to_hash_witness([0, 4], [[h(10), h(20), h(30)], [h(11), h(21), h(31)]], 3)
    # level 0: siblings 1 and 5; level 1: parents 0 and 2 need 1 and 3; level 2: 0 and 1 pair up
    assert w == [tuple(h(10)), tuple(h(11)), tuple(h(20)), tuple(h(21))]


@pytest.mark.parametrize("positions,vals,want", [
    ([0, 4, 8], [[1, 0, 0, 0], [5, 0, 0, 0], [9, 0, 0, 0]], [(1, 0, 0, 0), (5, 0, 0, 0), (9, 0, 0, 0)]),
    ([0, 1, 4, 5], [[1, 0, 0, 0], [0, 0, 0, 0], [5, 0, 0, 0], [4, 0, 0, 0]], []),
    ([0, 1, 4], [[1, 0, 0, 0], [0, 0, 0, 0], [5, 0, 0, 0]], [(5, 0, 0, 0)]),
    ([3, 3], [[2, 0, 0, 0], [2, 0, 0, 0]], [(2, 0, 0, 0)]),
])
def test_dedup_fri_witness(positions, vals, want):
    assert se.dedup_fri_witness(positions, vals) == want


# ------------------------------------------------- independent reconstruction
def b2_words(words):
    return hashlib.blake2s(b"".join(int(w).to_bytes(4, "little") for w in words)).digest()


def py_root(log_h, positions, cols, witness):
    """the stwo Merkle verifier walk, leaves = Blake2s of the row's LE words"""
    layer = {}
    for i, p in enumerate(positions):
        layer.setdefault(p, b2_words([c[i] for c in cols]))
    wit = iter(bytes(w) for w in witness)
    for _ in range(log_h):
        nxt = {}
        keys = sorted(layer)
        i = 0
        while i < len(keys):
            p = keys[i]
            if i + 1 < len(keys) and keys[i + 1] == p ^ 1:
                l, r = layer[p], layer[p ^ 1]
                i += 2
            else:
                w = next(wit)
                l, r = (w, layer[p]) if p & 1 else (layer[p], w)
                i += 1
            nxt[p >> 1] = hashlib.blake2s(l + r).digest()
        layer = nxt
    assert next(wit, None) is None, "witness not consumed"
    assert len(layer) == 1
    return next(iter(layer.values()))


def tree(rng, log_h, n_cols):
    rows = [[rng.randrange(P) for _ in range(n_cols)] for _ in range(1 << log_h)]
    layers = [[b2_words(r) for r in rows]]
    while len(layers[-1]) > 1:
        lv = layers[-1]
        layers.append([hashlib.blake2s(lv[2 * i] + lv[2 * i + 1]).digest() for i in range(len(lv) // 2)])
    return rows, layers


def path_words(layers, pos):
    out = []
    for lv in layers[:-1]:
        out.append(list(np.frombuffer(lv[pos ^ 1], np.uint32)))
        pos >>= 1
    return out


@pytest.mark.parametrize("seed,log_h,n_cols,n_pos", [(1, 5, 1, 6), (2, 8, 4, 40), (3, 6, 17, 64), (4, 3, 2, 1)])
def test_witness_rebuilds_the_root(seed, log_h, n_cols, n_pos):
    rng = random.Random(seed)
    rows, layers = tree(rng, log_h, n_cols)
    positions = [rng.randrange(1 << log_h) for _ in range(n_pos)]
    wit = se.auth_paths_to_hash_witness(positions, [path_words(layers, p) for p in positions], log_h)
    cols = [[rows[p][c] for p in positions] for c in range(n_cols)]
    root = layers[-1][0]
    wbytes = [np.array(w, np.uint32).tobytes() for w in wit]
    assert py_root(log_h, positions, cols, wbytes) == root
    assert se.reconstruct_root_from_witness(log_h, positions, cols, wit) == tuple(np.frombuffer(root, np.uint32))
    if wit:
        with pytest.raises(VortexError, match="too short"):
            se.reconstruct_root_from_witness(log_h, positions, cols, wit[:-1])
    with pytest.raises(VortexError, match="too long"):
        se.reconstruct_root_from_witness(log_h, positions, cols, wit + [tuple(h(7))])


# ------------------------------------------------------- last layer LinePoly
def double_x(x):
    return pyref.qm_sub(pyref.qm_scale(pyref.qm_mul(x, x), 2), (1, 0, 0, 0))


def line_eval(coeffs_brt, x):
    """stwo LinePoly::eval_at_point: fold the bit-reversed coefficients
    with the doublings x, 2x^2 - 1, ... (the first halving takes x)"""
    doublings = []
    for _ in range(len(coeffs_brt).bit_length() - 1):
        doublings.append(x)
        x = double_x(x)

    def fold(vals, fs):
        if len(vals) == 1:
            return vals[0]
        half = len(vals) // 2
        return pyref.qm_add(fold(vals[:half], fs[1:]), pyref.qm_mul(fold(vals[half:], fs[1:]), fs[0]))

    return fold(list(coeffs_brt), doublings)


def last_layer_points(log_n):
    c = pyref.half_odds(log_n)
    return [(pyref.coset_at(c, pyref.bitrev(i, log_n))[0], 0, 0, 0) for i in range(1 << log_n)]


@pytest.mark.parametrize("log_n", [1, 3, 5])
def test_last_layer_coeffs_invert_linepoly_evaluation(log_n):
    rng = random.Random(log_n)
    coeffs = [tuple(rng.randrange(P) for _ in range(4)) for _ in range(1 << log_n)]
    evals = [line_eval(coeffs, x) for x in last_layer_points(log_n)]
    assert se.last_layer_poly_coeffs(evals) == coeffs


# ------------------------------------------------------------ whole proofs
def proofs():
    return {
        "fib_v1": (po.prove(3, 5, 6, 2, 80, 0), 2),
        "fib_v2_pow": (po.prove(1, 1, 9, 2, 40, 6), 2),
        "fib_blowup3": (po.prove(2, 7, 7, 3, 24, 0), 3),
        "d2_v3": (po.d2_prove(5, 6, 34, 35, 0xD2A1, 2, 80, 0), 2),
        "d2_v3_ragged": (po.d2_prove(7, 5, 17, 4, 5, 2, 30, 4), 2),
        "batch_v4": (po.batch_prove(11, [3, 3, 5, 5, 5], 2, 16, 6), 2),
    }


@pytest.fixture(scope="module")
def exported():
    return {k: (se.proof_to_stwo(p, b), p, b) for k, (p, b) in proofs().items()}


CASES = ["fib_v1", "fib_v2_pow", "fib_blowup3", "d2_v3", "d2_v3_ragged", "batch_v4"]


@pytest.mark.parametrize("case", CASES)
def test_every_root_rebuilds_from_its_witness(exported, case):
    two = exported[case][0]
    ko = two["krak_openings"]
    entries = [(ko["trees"][t], two["decommitments"][t]["hash_witness"], two["commitments"][t])
               for t in range(len(two["commitments"]))]
    fl = two["fri_proof"]["first_layer"]
    entries.append((ko["first_layer"], fl["decommitment"]["hash_witness"], fl["commitment"]))
    for k, lay in enumerate(two["fri_proof"]["inner_layers"]):
        entries.append((ko["inner_layers"][k], lay["decommitment"]["hash_witness"], lay["commitment"]))
    for op, wit, commit in entries:
        assert py_root(op["log_size"], op["positions"], op["values"], wit) == bytes(commit)
    se.verify_all_merkle_witnesses(two)
    # the quotient tree is the FRI first layer
    assert fl["commitment"] == two["commitments"][-1]


@pytest.mark.parametrize("case", CASES)
def test_layout_and_queried_values(exported, case):
    two, proof, b = exported[case]
    cfg = two["config"]
    nq = cfg["n_queries"]
    assert len(two["query_indices"]) == nq and cfg["log_blowup_factor"] == b
    assert cfg["log_last_layer_degree_bound"] == 3
    n_trees = len(two["commitments"])
    assert len(two["tree_meta"]) == len(two["decommitments"]) == len(two["queried_values"]) == n_trees
    assert all(v == [] for v in two["sampled_values"])
    le = two["tree_meta"][0]["log_eval_size"]
    assert two["tree_meta"][-1] == {"name": "quotient", "n_cols": 4, "log_eval_size": le}
    assert len(two["fri_proof"]["inner_layers"]) == le - 3   # folds down to 8 values
    ko = two["krak_openings"]
    for t, meta in enumerate(two["tree_meta"]):
        qv = two["queried_values"][t]
        assert len(qv) == meta["n_cols"] and all(len(c) == nq for c in qv)
        op = ko["trees"][t]
        at = {p: i for i, p in enumerate(op["positions"])}
        for qi, q in enumerate(two["query_indices"]):
            assert [c[qi] for c in qv] == [c[at[q]] for c in op["values"]]
    # the trace openings of a query pair it with its partner row
    op = ko["trees"][0]
    part = (lambda q: (q + 1) % (1 << le)) if case.startswith("d2") else (lambda q: q ^ 1)
    assert set(op["positions"]) == set(two["query_indices"]) | {part(q) for q in two["query_indices"]}
    assert se.proof_to_json(proof, b) == json.dumps(two, separators=(",", ":"))


def unpaired(positions):
    s = set(positions)
    return sorted(p for p in s if p ^ 1 not in s)


@pytest.mark.parametrize("case", CASES)
def test_fri_witness_is_the_unpaired_siblings(exported, case):
    two = exported[case][0]
    ko = two["krak_openings"]
    layers = [(two["fri_proof"]["first_layer"], ko["first_layer"], 0)]
    layers += [(lay, ko["inner_layers"][k], k + 1) for k, lay in enumerate(two["fri_proof"]["inner_layers"])]
    for lay, op, shift in layers:
        pos = [q >> shift for q in two["query_indices"]]
        at = {p: i for i, p in enumerate(op["positions"])}
        want = [[c[at[p ^ 1]] for c in op["values"]] for p in unpaired(pos)]
        assert lay["fri_witness"] == want
        # whole fold pairs are opened
        assert sorted(op["positions"]) == sorted({p & ~1 for p in pos} | {p | 1 for p in pos})


@pytest.mark.parametrize("case", CASES)
def test_last_layer_polynomial(exported, case):
    two, _, b = exported[case]
    coeffs = [tuple(c) for c in two["fri_proof"]["last_layer_poly"]]
    evals = [tuple(e) for e in two["krak_openings"]["last_layer_evaluations"]]
    assert len(coeffs) == 8
    assert [line_eval(coeffs, x) for x in last_layer_points(3)] == evals


def test_tampered_witness_is_rejected(exported):
    two = json.loads(json.dumps(exported["fib_v1"][0]))
    w = two["decommitments"][0]["hash_witness"]
    w[len(w) // 2][0] ^= 1
    with pytest.raises(VortexError, match="tree 0: root mismatch"):
        se.verify_all_merkle_witnesses(two)


def test_bad_bytes_are_rejected():
    proof = po.prove(3, 5, 6, 2, 16, 0)
    for bad in (proof[:-4], proof + b"\0\0\0\0", b"XXXX" + proof[4:], proof[:4] + b"\x09\0\0\0" + proof[8:]):
        with pytest.raises(VortexError):
            se.proof_to_json(bad, 2)
