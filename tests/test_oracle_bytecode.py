"""The bytecode constraint evaluator in the oracle (or_bytecode_constraint_eval,
a restatement of cuda/constraint_eval.cu:97-407) and the product's host pass
(bc_compile via vx_bytecode_slots, host code).

Pinned by: the reference's three GPU tests of the kernel
(crates/cuda-backend/src/tests.rs:620-900: flag*(1-flag) on [0,1,0,1],
t-5 on [3,5,7,5], a+b-c on three columns); the row offset against a walk on
the circle (tests/golden/pyref_bytecode.py: stwo's CircleDomain points and
the trace-domain generator); random programs over every opcode, with
registers reused and redefined and row offsets, against the direct
evaluation of their op lists (no decoding)."""
import ctypes as C
import random

import numpy as np
import pytest

import pyoracle as po
import pyref_bytecode as pb
from vortexstark_amd import _lib
from vortexstark_amd._lib import VortexError

P = po.P


def one():
    return np.array([1, 0, 0, 0], np.uint32)


def test_reference_vector_zero_constraints():
    words = [(0x03 << 24) | 0, 0, 0, (0x01 << 24) | 1, 1, (0x11 << 24) | 2, (1 << 16) | 0, (0x12 << 24) | 3,
             (0 << 16) | 2, (0x40 << 24) | 3]
    out = po.bytecode_eval(words, [np.array([0, 1, 0, 1], np.uint32)], 4, 4, one(), [1])
    assert all(not np.any(o) for o in out)


def test_reference_vector_nonzero_constraints():
    words = [(0x03 << 24) | 0, 0, 0, (0x14 << 24) | 1, 0 << 16, P - 5, (0x40 << 24) | 1]
    out = po.bytecode_eval(words, [np.array([3, 5, 7, 5], np.uint32)], 4, 4, one(), [1])
    assert [int(x) for x in out[0]] == [P - 2, 0, 2, 0]
    assert not np.any(out[1]) and not np.any(out[2]) and not np.any(out[3])


def test_reference_vector_via_encoder():
    p = pb.Prog()
    p.load_trace(0, 0, 0)
    p.load_trace(1, 1, 0)
    p.add(2, 0, 1)
    p.load_trace(3, 2, 0)
    p.sub(4, 2, 3)
    p.add_constraint(4)
    cols = [np.array(c, np.uint32) for c in ([1, 2, 3, 4], [10, 20, 30, 40], [11, 22, 33, 44])]
    out = po.bytecode_eval(p.encode(), cols, 4, 4, one(), [1])
    assert all(not np.any(o) for o in out)  # a + b - c = 0 on every row


@pytest.mark.parametrize("e", [2, 3, 5, 8])
def test_offset_row_is_a_step_on_the_circle(e):
    for d in range(1, e):
        for off in (-5, -2, -1, 1, 2, 3, 7, (1 << d) + 1, -(1 << d) - 3, 1000):
            for i in range(1 << e):
                assert po.bc_offset_row(i, d, e, off) == pb.neighbor_row(i, d, e, off), (d, e, off, i)


@pytest.mark.parametrize("seed,n_rows,log_expand", [(1, 8, 1), (2, 16, 2), (3, 32, 1), (4, 4, 0), (5, 64, 3)])
def test_random_programs_match_direct_evaluation(seed, n_rows, log_expand):
    rng = random.Random(seed)
    trace_n = n_rows >> log_expand
    offs = (0,) if log_expand == 0 else (0, 0, 1, -1, 2)
    n_cols = 5
    prog = pb.random_program(rng, 120, n_cols, n_regs=30, offsets=offs)
    cols = [np.array([rng.randrange(P) for _ in range(n_rows)], np.uint32) for _ in range(n_cols)]
    cols[3] = None  # a null column reads as zero
    coeffs = [pb.rand_q(rng) for _ in range(prog.n_constraints())]
    denom = [rng.randrange(1, P) for _ in range(1 << log_expand)]
    want = prog.run(cols, n_rows, trace_n, coeffs, denom)
    got = po.bytecode_eval(prog.encode(), cols, n_rows, trace_n, np.array(coeffs, np.uint32).ravel(), denom)
    for c in range(4):
        assert [int(x) for x in got[c]] == want[c], c


def test_accumulates_into_existing_values():
    rng = random.Random(9)
    prog = pb.random_program(rng, 40, 2, offsets=(0,))
    cols = [np.array([rng.randrange(P) for _ in range(8)], np.uint32) for _ in range(2)]
    coeffs = np.array([pb.rand_q(rng) for _ in range(prog.n_constraints())], np.uint32).ravel()
    base = [np.array([rng.randrange(P) for _ in range(8)], np.uint32) for _ in range(4)]
    fresh = po.bytecode_eval(prog.encode(), cols, 8, 8, coeffs, [7])
    got = po.bytecode_eval(prog.encode(), cols, 8, 8, coeffs, [7], acc=base)
    for c in range(4):
        assert [int(x) for x in got[c]] == [(int(a) + int(b)) % P for a, b in zip(base[c], fresh[c])]


def test_oracle_rejects_malformed_programs():
    col = [np.zeros(4, np.uint32)]
    for words in ([(0x77 << 24) | 0], [(0x14 << 24) | 1, 0], [(0x03 << 24) | 0, 5, 0]):
        with pytest.raises(ValueError):
            po.bytecode_eval(words, col, 4, 4, one(), [1])


def slots(words, n_cols):
    w = np.array(words, np.uint32)
    m, q, k = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _lib.call("vx_bytecode_slots", w.ctypes.data if w.size else None, w.size, n_cols, C.byref(m), C.byref(q),
              C.byref(k))
    return m.value, q.value, k.value


def test_host_pass_renames_to_live_values():
    # a chain: every value is read once by the next instruction -> one slot
    p = pb.Prog()
    p.load_trace(0, 0, 0)
    for r in range(1, 200):
        p.add_const(r, r - 1, 3)
    p.add_constraint(199)
    assert slots(p.encode(), 1) == (1, 0, 1)
    # 100 loads of 3 columns, then summed: one slot per distinct (column,
    # offset), filled by the prologue, plus the running sum
    p = pb.Prog()
    for r in range(100):
        p.load_trace(r, r % 3, 0)
    for r in range(1, 100):
        p.add(0, 0, r)
    p.widen(500, 0)
    p.wide_mul(501, 500, 500)
    p.add_constraint(501)
    p.add_constraint(500)
    assert slots(p.encode(), 3) == (4, 2, 2)  # 500 stays live until the last constraint
    # 100 computed values live at once
    p = pb.Prog()
    p.load_trace(1000, 0, 0)
    for r in range(100):
        p.mul_const(r, 1000, r + 1)
    for r in range(1, 100):
        p.add(0, 0, r)
    p.add_constraint(0)
    assert slots(p.encode(), 1) == (100, 0, 1)


def test_host_pass_rejects_what_the_reference_would_misread():
    bad = {
        "unknown opcode": [(0x77 << 24) | 0],
        "truncated": [(0x14 << 24) | 1, 0],
        "column out of range": [(0x03 << 24) | 0, 5, 0],
        "read before written": [(0x10 << 24) | 2, (0 << 16) | 1],
        "non-canonical constant": [(0x01 << 24) | 0, P],
    }
    for name, words in bad.items():
        with pytest.raises(VortexError):
            slots(words, 2)
    # a redefinition after the last read is fine
    p = pb.Prog()
    p.load_const(0, 5)
    p.add(1, 0, 0)
    p.load_const(0, 7)
    p.mul(1, 1, 0)
    p.add_constraint(1)
    assert slots(p.encode(), 1)[2] == 1


def test_compiled_tier_sources_compile():
    # the straight-line source of random programs (every opcode, offsets,
    # redefinitions) compiles for gfx950 with hipRTC (no device needed)
    for seed in range(3):
        rng = random.Random(100 + seed)
        prog = pb.random_program(rng, 200, 6, n_regs=20, offsets=(0, 1, -1, 2))
        w = np.array(prog.encode(), np.uint32)
        n = C.c_size_t()
        _lib.call("vx_bytecode_jit_compile", w.ctypes.data, w.size, 6, 12, 10, C.byref(n))
        assert n.value > 0
    with pytest.raises(VortexError):  # checked like an evaluation
        w = np.array([(0x10 << 24) | 2, (0 << 16) | 1], np.uint32)
        _lib.call("vx_bytecode_jit_compile", w.ctypes.data, w.size, 1, 4, 4, C.byref(n))
