"""Merkle-shape autotuner: proofs made with tuned shapes equal the oracle's
and those made with the built-in shapes; the cache file round-trips and
ignores shapes this build does not accept for a key."""
import hashlib

import pytest

import vortexstark_amd as vx

pytestmark = pytest.mark.gpu

LEAF_COLS, LEAF_LIFTED = 0, 5


def valid_shape(e):
    # one leaf per lane with at most 256-leaf tiles, or 2^(tile - 8) leaves per lane
    if e["lane_log"] == 0:
        return e["tile_log"] == min(e["log_n"], 8)
    return e["tile_log"] - e["lane_log"] == 8 and e["tile_log"] <= e["log_n"]


@pytest.fixture(autouse=True)
def restore_tuner():
    yield
    vx.tuner(True, None)


def test_tuned_and_builtin_shapes_prove_the_same(golden):
    g = next(x for x in golden["proofs_c"] if (x["log_n"], x["blowup_bits"], x["pow_bits"]) == (20, 2, 0))
    vx.tuner(enabled=False)
    p = vx.Prover(20, graphs=False)
    builtin = p.prove(g["a"], g["b"])
    p.close()
    vx.tuner(enabled=True)
    for graphs in (False, True):
        q = vx.Prover(20, graphs=graphs)
        tuned = q.prove(g["a"], g["b"])
        assert tuned == q.prove(g["a"], g["b"])
        q.close()
        assert tuned == builtin
    assert hashlib.sha256(builtin).hexdigest() == g["sha256"]
    ents = vx.tuner_entries()
    # the trace tree of this class (2^22 one-column leaves) was measured
    assert any(e["leaf_kind"] == LEAF_COLS and e["n_cols"] == 1 and e["log_n"] == 22 for e in ents), ents
    assert all(valid_shape(e) for e in ents), ents


def test_tuner_cache_file(tmp_path):
    f = tmp_path / "tune.txt"
    # kind n_cols n_cons lift_n log_n store_lo tile_log lpt: a valid entry
    # for a key no test commits, and one with a shape no key accepts
    f.write_text("0 3 0 0 29 4 10 2\n0 3 0 0 28 4 9 1\n")
    vx.tuner(True, str(f))
    ents = vx.tuner_entries()
    assert any((e["n_cols"], e["log_n"], e["tile_log"], e["lane_log"]) == (3, 29, 10, 2) for e in ents)
    assert not any((e["n_cols"], e["log_n"]) == (3, 28) for e in ents)
    # a lifted tree of a shape no other test uses is measured and appended
    b = vx.Batch([14, 15, 15, 16, 17], 2, True)
    b.fill_random(5)
    r1 = b.run()
    assert b.run() == r1
    b.close()
    lines = f.read_text().split("\n")
    assert any(ln.split()[:5] == ["5", "0", "0", "5", "19"] for ln in lines if ln), lines
