"""stwo wire format for the KRAK proofs of this library (src/stwo_export.rs).

`proof_to_json` is the cairo_proof_to_json counterpart (src/stwo_export.rs:
673-676) for KRAK v1..v4 bytes; the helpers are the reference's public
witness functions. All of it is host code in libvortex_hip.so
(csrc/export.cpp); nothing here needs a GPU."""
import ctypes as C
import json

import numpy as np

from . import _lib
from ._lib import VortexError, load

__all__ = ["proof_to_json", "proof_to_stwo", "auth_paths_to_hash_witness", "dedup_fri_witness",
           "reconstruct_root_from_witness", "last_layer_poly_coeffs", "verify_all_merkle_witnesses"]


def _ptr(a):
    return a.ctypes.data if a.size else None


def proof_to_json(proof: bytes, blowup_bits=2) -> str:
    """KRAK proof bytes -> stwo-shaped JSON text (plus the krak_openings
    extension: the leaves each hash witness proves)."""
    n = C.c_size_t()
    _lib.call("vx_export_stwo_json", proof, len(proof), blowup_bits, None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    _lib.call("vx_export_stwo_json", proof, len(proof), blowup_bits, buf, n.value + 1, C.byref(n))
    return buf.value.decode()


def proof_to_stwo(proof: bytes, blowup_bits=2) -> dict:
    return json.loads(proof_to_json(proof, blowup_bits))


def auth_paths_to_hash_witness(positions, auth_paths, tree_height):
    """src/stwo_export.rs:262-320. auth_paths[i] = tree_height hashes of 8
    words (leaf level first). Returns a list of 8-word tuples."""
    pos = np.asarray(positions, np.uint64)
    paths = np.zeros((len(pos), tree_height, 8), np.uint32)
    for i, p in enumerate(auth_paths):
        paths[i] = np.asarray(p, np.uint32).reshape(tree_height, 8)
    n = C.c_size_t()
    _lib.call("vx_stwo_hash_witness", _ptr(pos), len(pos), _ptr(paths), tree_height, None, 0, C.byref(n))
    out = np.zeros((max(n.value, 1), 8), np.uint32)
    _lib.call("vx_stwo_hash_witness", _ptr(pos), len(pos), _ptr(paths), tree_height, out.ctypes.data, n.value,
              C.byref(n))
    return [tuple(int(x) for x in h) for h in out[:n.value]]


def dedup_fri_witness(positions, sibling_values):
    """src/stwo_export.rs:174-206: sibling_values[i] (QM31, 4 words) is the
    value at positions[i] ^ 1."""
    pos = np.asarray(positions, np.uint64)
    sib = np.asarray(sibling_values, np.uint32).reshape(-1, 4)
    n = C.c_size_t()
    _lib.call("vx_stwo_fri_witness", _ptr(pos), len(pos), _ptr(sib), None, 0, C.byref(n))
    out = np.zeros((max(n.value, 1), 4), np.uint32)
    _lib.call("vx_stwo_fri_witness", _ptr(pos), len(pos), _ptr(sib), out.ctypes.data, n.value, C.byref(n))
    return [tuple(int(x) for x in v) for v in out[:n.value]]


def reconstruct_root_from_witness(log_height, positions, col_values, hash_witness):
    """src/stwo_export.rs:697-780: the root (8 words) from the leaf values
    (col_values[col][i] at positions[i]) and the hash witness (8-word
    hashes, or 32-byte TwoHash lists). Raises VortexError on a bad witness."""
    pos = np.asarray(positions, np.uint64)
    cols = np.ascontiguousarray(np.asarray(col_values, np.uint32).reshape(-1, len(pos)))
    wit = np.zeros((max(len(hash_witness), 1), 8), np.uint32)
    for i, h in enumerate(hash_witness):
        h = list(h)
        wit[i] = np.frombuffer(bytes(h), np.uint32) if len(h) == 32 else h
    root = np.zeros(8, np.uint32)
    err = C.create_string_buffer(256)
    rc = load().vx_stwo_reconstruct_root(log_height, _ptr(pos), len(pos), _ptr(cols), cols.shape[0],
                                          wit.ctypes.data, len(hash_witness), root.ctypes.data, err, 256)
    if rc != 0:
        raise VortexError("vx_stwo_reconstruct_root", rc, err.value.decode())
    return tuple(int(x) for x in root)


def last_layer_poly_coeffs(evals):
    """src/fri.rs:294-335 + the bit-reversal of src/cairo_air/prover.rs:
    2896-2900: bit-reversed last-layer evaluations -> bit-reversed LinePoly
    coefficients."""
    ev = np.ascontiguousarray(np.asarray(evals, np.uint32).reshape(-1, 4))
    out = np.zeros_like(ev)
    _lib.call("vx_stwo_last_layer_coeffs", ev.ctypes.data, ev.shape[0], out.ctypes.data)
    return [tuple(int(x) for x in v) for v in out]


def _hash_words(h):
    return tuple(int(x) for x in np.frombuffer(bytes(h), np.uint32))


def verify_all_merkle_witnesses(two: dict):
    """src/stwo_export.rs:796-880 over an exported document: every tree
    root, the FRI first layer and each inner layer rebuilt from its
    krak_openings leaves and hash witness must equal its commitment."""
    ko = two["krak_openings"]
    cases = [(f"tree {t}", ko["trees"][t], two["decommitments"][t]["hash_witness"], two["commitments"][t])
             for t in range(len(two["commitments"]))]
    fl = two["fri_proof"]["first_layer"]
    cases.append(("fri first layer", ko["first_layer"], fl["decommitment"]["hash_witness"], fl["commitment"]))
    for k, lay in enumerate(two["fri_proof"]["inner_layers"]):
        cases.append((f"fri layer {k}", ko["inner_layers"][k], lay["decommitment"]["hash_witness"],
                      lay["commitment"]))
    for name, op, wit, commit in cases:
        root = reconstruct_root_from_witness(op["log_size"], op["positions"], op["values"], wit)
        if root != _hash_words(commit):
            raise VortexError("verify_all_merkle_witnesses", 4, f"{name}: root mismatch")
