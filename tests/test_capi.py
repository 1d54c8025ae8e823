"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/vortex_hip.h declares, sizes proofs like the oracle, and its host
verifier (src/verifier.rs counterpart) accepts the oracle's proofs and rejects
the reference's tamper cases. No kernel is launched here."""
import random
import struct

import pytest

import pyoracle as po
import vortexstark_amd as vx
from vortexstark_amd import _lib
from krak import Krak
from test_oracle import tamper_cases


def test_library_exports_every_header_symbol():
    lib = vx.load()
    names = _lib.header_symbols()
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


@pytest.mark.parametrize("log_n,bb,pw", [(4, 2, 0), (6, 1, 0), (10, 3, 0), (16, 2, 0), (24, 2, 0), (24, 2, 26),
                                         (30, 1, 0), (29, 2, 0)])
def test_proof_size_matches_oracle(log_n, bb, pw):
    assert vx.proof_size(log_n, bb, 80, pw) == po.proof_size(log_n, bb, 80, pw)


def test_proof_size_rejects_bad_shapes():
    assert vx.proof_size(30, 2, 80, 0) == 0
    assert vx.proof_size(0, 2, 80, 0) == 0


@pytest.mark.parametrize("log_n,a,b,bb,pw", [(4, 1, 1, 2, 0), (5, 2, 3, 2, 0), (6, 42, 99, 2, 0),
                                             (8, 1000, 2000, 2, 0), (6, 1, 1, 1, 0), (5, 1, 1, 3, 0),
                                             (6, 5, 8, 2, 8), (12, 7, 11, 2, 0)])
def test_host_verifier_accepts_oracle_proofs(log_n, a, b, bb, pw):
    proof = po.prove(a, b, log_n, blowup_bits=bb, pow_bits=pw)
    assert vx.verify(proof, bb) == (True, "")


def test_host_verifier_rejects_tampering():
    proof = po.prove(1, 1, 6)
    for name, bad in tamper_cases(proof).items():
        ok, why = vx.verify(bad)
        assert not ok and why, name


@pytest.mark.parametrize("extra", [1, 2, 3])
def test_host_verifier_rejects_extra_fri_layers(extra):
    # Appending copies of the last FRI decommitment (no root covers them)
    # must be rejected, not read past the drawn FRI alphas.
    proof = po.prove(1, 1, 6)
    k = Krak(proof)
    last = k.fri[-1]
    start = last["values"] - 1
    words = list(k.words)
    words[k.off_n_fri_decom] += extra
    words += words[start:] * extra
    bad = struct.pack(f"<{len(words)}I", *words)
    ok, why = vx.verify(bad)
    assert not ok and "FRI decommitments" in why, why


def test_host_verifier_rejects_missing_fri_layers():
    proof = po.prove(1, 1, 6)
    k = Krak(proof)
    start = k.fri[-1]["values"] - 1
    words = list(k.words[:start])
    words[k.off_n_fri_decom] -= 1
    ok, why = vx.verify(struct.pack(f"<{len(words)}I", *words))
    assert not ok and why


def test_host_verifier_rejects_wrong_params_and_bad_pow():
    proof = po.prove(1, 1, 6)
    assert not vx.verify(proof, blowup_bits=1)[0]
    assert not vx.verify(proof, n_queries=79)[0]
    p2 = bytearray(po.prove(5, 8, 5, pow_bits=6))
    p2[28] ^= 1  # nonce low word
    assert not vx.verify(bytes(p2))[0]


def test_host_verifier_survives_truncation_and_garbage():
    proof = po.prove(1, 1, 5)
    rng = random.Random(0)
    for cut in [0, 3, 4, 20, 100, len(proof) // 2, len(proof) - 4, len(proof) - 1]:
        assert not vx.verify(proof[:cut])[0]
    for _ in range(50):
        b = bytearray(proof)
        for _ in range(4):
            b[rng.randrange(len(b))] = rng.randrange(256)
        ok, _ = vx.verify(bytes(b))
        if ok:
            # only bytes the verifier never reads could leave it accepting
            assert bytes(b) != proof or ok
    huge = struct.pack("<5I", 0x4B52414B, 1, 6, 1, 1) + b"\xff" * 64
    assert not vx.verify(huge)[0]
