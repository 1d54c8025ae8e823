"""Pins the CPU oracle (oracle/vx_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong index or sign, a transposed operand) fails: reference literal
KATs, hashlib Blake2s, closed forms derived by hand from the reference's layer
definitions, algebraic invariants, and byte equality with an independent
pure-Python restatement (tests/golden/pyref.py) at small sizes.
"""
import hashlib
import random
import struct

import numpy as np
import pytest

import pyoracle as po
import pyref
from krak import Krak

P = po.P
L = po.lib()


# ---------------------------------------------------------------- field
def test_m31_reference_kats():
    # src/field/m31.rs:145-178
    assert L.or_m31_add(5, 3) == 8
    assert L.or_m31_add(P - 1, 1) == 0
    assert L.or_m31_add(P - 1, 2) == 1
    assert L.or_m31_sub(5, 3) == 2
    assert L.or_m31_sub(0, 1) == P - 1
    assert L.or_m31_mul(5, 3) == 15
    assert L.or_m31_mul(1 << 16, 1 << 16) == (1 << 32) % P
    assert L.or_m31_mul(L.or_m31_inv(12345), 12345) == 1
    assert L.or_m31_inv(0) == 0  # Fermat inverse convention, m31.rs:32-35


def test_m31_matches_bigint():
    rng = random.Random(1)
    for _ in range(2000):
        a, b = rng.randrange(P), rng.randrange(P)
        assert L.or_m31_mul(a, b) == a * b % P
        assert L.or_m31_add(a, b) == (a + b) % P
        assert L.or_m31_sub(a, b) == (a - b) % P
    for a in [1, 2, P - 1, P - 2, 1 << 30]:
        assert L.or_m31_inv(a) == pow(a, P - 2, P)


def test_cm31_reference_kat():
    out = np.zeros(2, np.uint32)
    L.or_cm31_mul(po.u32([3, 4]), po.u32([5, 6]), out)  # cm31.rs:117-137
    assert list(out) == [P - 9, 38]


def test_qm31_extension_relations():
    i = [0, 1, 0, 0]
    u = [0, 0, 1, 0]
    assert list(po.qm31_mul(i, i)) == [P - 1, 0, 0, 0]      # i^2 = -1
    assert list(po.qm31_mul(u, u)) == [2, 1, 0, 0]          # u^2 = 2 + i (qm31.rs:1-2)
    assert list(po.qm31_mul(u, i)) == [0, 0, 0, 1]          # word order [a.re, a.im, b.re, b.im]
    a = [111, 222, 333, 444]                               # qm31.rs:186-192
    inv = np.zeros(4, np.uint32)
    L.or_qm31_inv(po.u32(a), inv)
    assert list(po.qm31_mul(a, inv)) == [1, 0, 0, 0]


def test_qm31_field_axioms():
    rng = random.Random(2)
    for _ in range(200):
        a, b, c = ([rng.randrange(P) for _ in range(4)] for _ in range(3))
        ab_c = po.qm31_mul(po.qm31_mul(a, b), c)
        a_bc = po.qm31_mul(a, po.qm31_mul(b, c))
        assert list(ab_c) == list(a_bc)
        assert list(po.qm31_mul(a, b)) == list(po.qm31_mul(b, a))
        assert list(po.qm31_mul(a, b)) == list(pyref.qm_mul(a, b))


# --------------------------------------------------------------- circle
def test_circle_generator_and_order():
    g = po.coset(po.SUBGROUP, 31)  # step = G
    gx, gy = int(g[2]), int(g[3])
    assert (gx, gy) == (2, 1268011823)
    assert (gx * gx + gy * gy) % P == 1  # circle.rs:589-595
    # G^(2^31) = identity but G^(2^30) != identity (order exactly 2^31)
    assert po.coset_at(po.coset(po.SUBGROUP, 31), 1 << 31) == (1, 0)
    assert po.coset_at(po.coset(po.SUBGROUP, 31), 1 << 30) != (1, 0)
    sg = po.coset(po.SUBGROUP, 2)
    assert po.coset_at(sg, 4) == (1, 0) and po.coset_at(sg, 2) != (1, 0)


def test_coset_at_matches_iteration():
    for kind in (po.SUBGROUP, po.HALF_COSET, po.HALF_ODDS):
        c = po.coset(kind, 6)
        p = (int(c[0]), int(c[1]))
        for i in range(64):
            assert po.coset_at(c, i) == p
            p = pyref.pt_mul(p, (int(c[2]), int(c[3])))


def test_half_coset_is_a_coset_of_size_2L():
    c = po.coset(po.HALF_COSET, 5)
    pts = {po.coset_at(c, i) for i in range(32)}
    assert len(pts) == 32
    assert all((x * x + y * y) % P == 1 for x, y in pts)
    # half_coset(L) = odd multiples of G^(2^(30-L)): it is closed under
    # conjugation and is the complement of subgroup(L) in subgroup(L+1)
    assert {(x, (P - y) % P) for x, y in pts} == pts
    sub = {po.coset_at(po.coset(po.SUBGROUP, 5), i) for i in range(32)}
    assert not (pts & sub)
    assert pts | sub == {po.coset_at(po.coset(po.SUBGROUP, 6), i) for i in range(64)}


# -------------------------------------------------------------- blake2s
def test_blake2s_vs_hashlib(golden):
    for kat in golden["blake2s"]:
        data = bytes.fromhex(kat["in"])
        assert po.blake2s(data).hex() == kat["out"]
        assert hashlib.blake2s(data).hexdigest() == kat["out"]
        if len(data) <= 64:
            out = np.zeros(32, np.uint8)
            L.or_blake2s_single(data, len(data), out)
            assert out.tobytes().hex() == kat["out"]


def test_blake2s_single_block_truncates_to_64_bytes():
    # src/channel.rs:254-257: the single-block helper hashes min(len, 64) bytes
    data = bytes(range(100))
    out = np.zeros(32, np.uint8)
    L.or_blake2s_single(data, len(data), out)
    assert out.tobytes() == hashlib.blake2s(data[:64]).digest()


# -------------------------------------------------------------- channel
def test_channel_transcript(golden):
    ops = golden["channel"]
    ch = po.PyChannel()
    ch.mix_digest(ops[0]["digest"])
    assert ch.state.hex() == ops[0]["state"]
    for op in ops[1:4]:
        assert list(ch.draw_felt()) == op["out"]
    ch.mix_felts(ops[4]["felts"])
    assert ch.state.hex() == ops[4]["state"]
    assert ch.draw_number(ops[5]["bound"]) == ops[5]["out"]
    ch.mix_u64(ops[6]["v"])
    assert ch.state.hex() == ops[6]["state"]


def test_channel_squeeze_layout():
    # draw_number = LE u64 of Blake2s(state || n_draws || 0x00) mod bound
    ch = po.PyChannel()
    ch.mix_digest(list(range(8)))
    st = ch.state
    for k in range(3):
        h = hashlib.blake2s(st + struct.pack("<I", k) + b"\x00").digest()
        assert ch.draw_number(1 << 40) == struct.unpack("<Q", h[:8])[0] % (1 << 40)


def test_pow_nonce_is_smallest(golden):
    for kat in golden["pow"]:
        ch = po.PyChannel()
        ch.mix_digest(kat["digest"])
        assert ch.grind(kat["bits"]) == kat["nonce"]
        assert ch.verify_pow_nonce(kat["bits"], kat["nonce"])
        assert not any(ch.verify_pow_nonce(kat["bits"], n) for n in range(kat["nonce"]))


# ------------------------------------------------------------- twiddles
KINDS = {"half_coset": po.HALF_COSET, "half_odds": po.HALF_ODDS, "subgroup": po.SUBGROUP}


def test_twiddle_tables_vs_closed_form(golden):
    for t in golden["twiddles"]:
        line, circ, offs, sizes = po.twiddles(po.coset(KINDS[t["kind"]], t["log"]))
        n = 1 << t["log"]
        assert list(circ[:n // 2]) == t["circle"]
        for l, layer in enumerate(t["line"]):
            assert int(sizes[l]) == len(layer)
            assert list(line[offs[l]:offs[l] + sizes[l]]) == layer


def test_batch_inverse_zero_convention():
    x = po.u32([0, 5, 0, 7, P - 1, 0])
    got = po.batch_inverse(x)
    assert list(got) == [0 if v == 0 else pow(int(v), P - 2, P) for v in x]


def test_fold_twiddles_vs_closed_form(golden):
    for t in golden["fold_twiddles"]:
        kind = po.HALF_COSET if t["kind"].startswith("half_coset") else po.HALF_ODDS
        got = po.fold_twiddles(po.coset(kind, t["log"]), t["kind"].endswith("_y"))
        assert list(got) == t["tw"]


# ------------------------------------------------------------------ NTT
def test_ntt_n4_closed_form():
    # cuda/circle_ntt.cu:248-275 worked by hand for N = 4:
    #   out = [c0+c1y0+c2x0+c3x0y0, c0-c1y0+c2x0-c3x0y0, c0+c1y1-c2x0-c3x0y1, c0-c1y1-c2x0+c3x0y1]
    c4 = po.coset(po.HALF_COSET, 2)
    (x0, y0), (_, y1) = po.coset_at(c4, 0), po.coset_at(c4, 1)
    rng = random.Random(3)
    for _ in range(20):
        c = [rng.randrange(P) for _ in range(4)]
        want = [(c[0] + c[1] * y0 + c[2] * x0 + c[3] * x0 * y0) % P,
                (c[0] - c[1] * y0 + c[2] * x0 - c[3] * x0 * y0) % P,
                (c[0] + c[1] * y1 - c[2] * x0 - c[3] * x0 * y1) % P,
                (c[0] - c[1] * y1 - c[2] * x0 + c[3] * x0 * y1) % P]
        assert list(po.evaluate(c, c4)) == want


def test_ntt_basis_and_random_vs_pyref(golden):
    for t in golden["ntt_basis"]:
        c = po.coset(po.HALF_COSET, t["log"])
        n = 1 << t["log"]
        for k in range(n):
            e = np.zeros(n, np.uint32)
            e[k] = 1
            assert list(po.evaluate(e, c)) == t["evaluate_unit"][k]
    for t in golden["ntt_random"]:
        c = po.coset(po.HALF_COSET, t["log"])
        assert list(po.evaluate(t["in"], c)) == t["evaluate"]
        assert list(po.interpolate(t["in"], c)) == t["interpolate"]


@pytest.mark.parametrize("log_n", [2, 3, 4, 7, 10, 12])
def test_ntt_roundtrip(log_n):
    # src/ntt.rs:386-449
    rng = np.random.default_rng(log_n)
    v = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    c = po.coset(po.HALF_COSET, log_n)
    assert np.array_equal(po.interpolate(po.evaluate(v, c), c), v)
    assert np.array_equal(po.evaluate(po.interpolate(v, c), c), v)


def test_ntt_layerwise_equals_evaluate():
    # src/prover.rs:2490-2574: evaluate == line layers high->low then circle layer
    log_n = 9
    c = po.coset(po.HALF_COSET, log_n)
    line, circ, offs, _ = po.twiddles(c)
    v = np.random.default_rng(9).integers(0, P, 1 << log_n, dtype=np.uint32)
    d = v.copy()
    for l in reversed(range(log_n - 1)):
        d = po.ntt_layer(d, line[offs[l]:], l + 1, True)
    d = po.ntt_layer(d, circ, 0, True)
    assert np.array_equal(d, po.evaluate(v, c))


def test_ntt_linearity():
    c = po.coset(po.HALF_COSET, 8)
    rng = np.random.default_rng(5)
    u, v = (rng.integers(0, P, 256, dtype=np.uint32) for _ in range(2))
    k = 987654321
    lhs = po.evaluate((u.astype(np.uint64) * k + v) % P, c)
    rhs = (po.evaluate(u, c).astype(np.uint64) * k + po.evaluate(v, c)) % P
    assert np.array_equal(lhs, rhs.astype(np.uint32))


def test_interpolate_scale_is_inverse_of_n():
    for log_n in range(1, 31):
        e = (30 * log_n) % 31
        assert (pow(2, e, P) * pow(2, log_n, P)) % P == 1


@pytest.mark.parametrize("log_n,blowup", [(4, 2), (6, 1), (7, 3)])
def test_lde_first_layers_replicate(log_n, blowup):
    # A.5: with zero padding, the top `blowup` evaluate layers copy the
    # coefficients into every block; the LDE equals the remaining layers run
    # on the tiled coefficients.
    le = log_n + blowup
    c = po.eval_domain(le)
    line, circ, offs, _ = po.twiddles(c)
    coef = np.random.default_rng(le).integers(0, P, 1 << log_n, dtype=np.uint32)
    padded = np.zeros(1 << le, np.uint32)
    padded[:1 << log_n] = coef
    d = padded.copy()
    for l in reversed(range(le - 1 - blowup, le - 1)):
        d = po.ntt_layer(d, line[offs[l]:], l + 1, True)
    assert np.array_equal(d, np.tile(coef, 1 << blowup))
    for l in reversed(range(le - 1 - blowup)):
        d = po.ntt_layer(d, line[offs[l]:], l + 1, True)
    d = po.ntt_layer(d, circ, 0, True)
    assert np.array_equal(d, po.evaluate(padded, c))


# ------------------------------------------------- quotient, folds, trace
def test_trace_is_fibonacci():
    t = po.fibonacci_trace(3, 7, 10)
    assert list(t[:6]) == [3, 7, 10, 17, 27, 44]
    assert all((int(t[i]) + int(t[i + 1])) % P == int(t[i + 2]) for i in range(len(t) - 2))


def test_quotient_vanishes_on_valid_trace():
    # On the trace itself the transition constraint holds except at the two
    # wrap-around rows, so alpha * c is zero there (cuda/constraint.cu:22-30).
    t = po.fibonacci_trace(1, 1, 8)
    q = po.quotient(t, [5, 6, 7, 8])
    nz = [i for i in range(len(t)) if any(int(col[i]) for col in q)]
    assert nz == [len(t) - 2, len(t) - 1]


def test_quotient_closed_form():
    rng = np.random.default_rng(11)
    t = rng.integers(0, P, 64, dtype=np.uint32)
    alpha = [int(x) for x in rng.integers(0, P, 4)]
    q = po.quotient(t, alpha)
    n = len(t)
    for i in range(n):
        c = (int(t[(i + 2) % n]) - int(t[(i + 1) % n]) - int(t[i])) % P
        assert [int(col[i]) for col in q] == [a * c % P for a in alpha]


def test_folds_closed_form():
    rng = np.random.default_rng(12)
    cols = [rng.integers(0, P, 32, dtype=np.uint32) for _ in range(4)]
    tw = rng.integers(0, P, 16, dtype=np.uint32)
    alpha = [int(x) for x in rng.integers(0, P, 4)]
    out = po.fold_line(cols, tw, alpha)
    for i in range(16):
        f0 = tuple(int(c[2 * i]) for c in cols)
        f1 = tuple(int(c[2 * i + 1]) for c in cols)
        assert tuple(int(o[i]) for o in out) == pyref.fold(f0, f1, tuple(alpha), int(tw[i]))
    # circle fold into a non-zero accumulator: dst * alpha^2 + fold
    dst = [rng.integers(0, P, 16, dtype=np.uint32) for _ in range(4)]
    a2 = [int(x) for x in po.qm31_mul(alpha, alpha)]
    got = po.fold_circle_into_line(dst, cols, tw, alpha, a2)
    for i in range(16):
        prev = tuple(int(d[i]) for d in dst)
        f0 = tuple(int(c[2 * i]) for c in cols)
        f1 = tuple(int(c[2 * i + 1]) for c in cols)
        want = pyref.qm_add(pyref.qm_mul(prev, tuple(a2)), pyref.fold(f0, f1, tuple(alpha), int(tw[i])))
        assert tuple(int(g[i]) for g in got) == want


# --------------------------------------------------------------- Merkle
@pytest.mark.parametrize("ncols,log_n", [(1, 0), (1, 5), (4, 3), (4, 8), (7, 4), (20, 2)])
def test_merkle_root_vs_hashlib(ncols, log_n):
    rng = np.random.default_rng(ncols * 100 + log_n)
    cols = [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(ncols)]
    rows = [[int(c[i]) for c in cols[:16]] for i in range(1 << log_n)]  # t = 4 * min(n_cols, 16)
    layers = pyref.merkle_layers(rows)
    assert list(po.merkle_root(cols)) == pyref.words(layers[-1][0])


def test_auth_path_verification():
    rng = np.random.default_rng(13)
    col = rng.integers(0, P, 64, dtype=np.uint32)
    layers = pyref.merkle_layers([[int(x)] for x in col])
    root = po.u32(pyref.words(layers[-1][0]))
    for idx in [0, 1, 17, 63]:
        path = po.u32(sum(pyref.auth_path(layers, idx), []))
        leaf = po.u32(pyref.words(layers[0][idx]))
        assert L.or_verify_auth_path(root, leaf, idx, path, 6) == 1
        assert L.or_verify_auth_path(root, leaf, idx ^ 1, path, 6) == 0


# ------------------------------------------------------- prove / verify
def test_proofs_match_independent_python(golden):
    for g in golden["proofs_py"]:
        data = po.prove(g["a"], g["b"], g["log_n"], blowup_bits=g["blowup_bits"], pow_bits=g["pow_bits"])
        assert len(data) == g["len"] == po.proof_size(g["log_n"], g["blowup_bits"], 80, g["pow_bits"])
        assert hashlib.sha256(data).hexdigest() == g["sha256"], g


def test_oracle_digests_stable(golden):
    for g in golden["proofs_c"]:
        if g["log_n"] > 16:
            continue
        data = po.prove(g["a"], g["b"], g["log_n"], blowup_bits=g["blowup_bits"], pow_bits=g["pow_bits"],
                        omp=True)
        assert hashlib.sha256(data).hexdigest() == g["sha256"], g


@pytest.mark.parametrize("log_n", [4, 5, 6, 8])
def test_verify_accepts_sizes(log_n):
    # src/verifier.rs:383-391
    assert po.verify(po.prove(1, 1, log_n)) == (True, "")


@pytest.mark.parametrize("a,b", [(1, 1), (2, 3), (42, 99), (1000, 2000)])
def test_verify_accepts_inputs(a, b):
    # src/verifier.rs:393-400
    assert po.verify(po.prove(a, b, 6))[0]


def tamper_cases(data):
    """The reference's tamper tests (src/verifier.rs:324-380) on KRAK bytes."""
    k = Krak(data)
    out = {}
    w = list(k.words)
    w[k.trace["values"]] ^= 1
    out["trace_value"] = w
    w = list(k.words)
    w[k.quotient["values"]] ^= 1
    out["quotient_value"] = w
    w = list(k.words)
    w[k.fri[0]["values"]] ^= 1
    out["fri_value"] = w
    w = list(k.words)
    w[k.off_trace_root] ^= 1
    out["commitment"] = w
    w = list(k.words)
    last = k.off_fri_roots[-1]
    w[k.off_fri_roots[0] - 1] -= 1
    del w[last:last + 8]
    out["dropped_fri_layer"] = w
    w = list(k.words)
    w[k.off_trace_root:k.off_trace_root + 8] = [0] * 8
    out["zero_commitment"] = w
    return {name: struct.pack(f"<{len(v)}I", *v) for name, v in out.items()}


def test_verify_rejects_tampering():
    data = po.prove(1, 1, 6)
    for name, bad in tamper_cases(data).items():
        ok, why = po.verify(bad)
        assert not ok, name


def test_prove_deterministic_and_input_sensitive():
    p1, p2 = po.prove(1, 1, 6), po.prove(1, 1, 6)
    assert p1 == p2
    k1, k2 = Krak(p1), Krak(po.prove(2, 3, 6))
    assert k1.root(k1.off_trace_root) != k2.root(k2.off_trace_root)


def test_pow_v2_nonce_minimal():
    data = po.prove(5, 8, 5, pow_bits=6)
    k = Krak(data)
    assert k.version == 2 and k.pow_bits == 6
    assert po.verify(data)[0]
    ch = pyref.Channel()
    # rebuild the transcript up to the PoW point from the proof's own fields
    ch.mix_digest(k.root(k.off_trace_root))
    ch.draw_felt()
    ch.mix_digest(k.root(k.off_quotient_root))
    for off in k.off_fri_roots:
        ch.draw_felt()
        ch.mix_digest(k.root(off))
    last = k.words[k.off_last:k.off_last + 32]
    ch.mix_felts([tuple(last[4 * i:4 * i + 4]) for i in range(8)])
    assert ch.pow_ok(6, k.nonce)
    assert not any(ch.pow_ok(6, n) for n in range(k.nonce))
