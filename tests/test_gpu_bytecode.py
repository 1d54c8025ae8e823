"""The bytecode constraint evaluator on the GPU (vx_bytecode_constraint_eval,
vx_warp_bytecode_constraint_eval through the C-ABI), compiled and
interpreted, against the oracle
(or_bytecode_constraint_eval, pinned by tests/test_oracle_bytecode.py): the
reference's three kernel tests, random programs over every opcode at
4..2^16 rows with row offsets, programs whose register file fits LDS and
programs that need the global workspace, accumulation into existing values,
null columns and rejected programs."""
import random

import numpy as np
import pytest

import pyoracle as po
import pyref_bytecode as pb
from gpu_util import dev, empty, host, sync
from vortexstark_amd._lib import VortexError, call

pytestmark = pytest.mark.gpu
P = po.P


def col_table(cols):
    """device columns (None stays a null pointer) and their device pointer array"""
    bufs = [None if c is None else dev(c) for c in cols]
    arr = np.array([0 if b is None else b.ptr for b in bufs], np.uint64).view(np.uint32)
    return bufs, dev(arr)


def gpu_eval(words, cols, n_rows, trace_n, coeff, denom, acc=None, entry="vx_bytecode_constraint_eval"):
    w = dev(np.array(words, np.uint32))
    bufs, table = col_table(cols)
    k = dev(np.asarray(coeff, np.uint32).ravel() if len(coeff) else np.zeros(4, np.uint32))
    dn = dev(np.asarray(denom, np.uint32))
    outs = [empty(n_rows) if acc is None else dev(acc[c]) for c in range(4)]
    log_expand = (n_rows // trace_n).bit_length() - 1
    call(entry, w.ptr, len(words), table.ptr, None, len(cols), n_rows, trace_n, k.ptr, dn.ptr, log_expand,
         *[o.ptr for o in outs], 64, None)
    sync()
    return [host(o) for o in outs]


def check(prog_words, cols, n_rows, trace_n, coeff, denom, acc=None, entry="vx_bytecode_constraint_eval"):
    """both tiers (compiled and interpreted) against the oracle"""
    want = po.bytecode_eval(prog_words, cols, n_rows, trace_n, np.asarray(coeff, np.uint32).ravel(), denom, acc)
    try:
        for jit in (1, 0):
            call("vx_bytecode_set_jit", jit)
            got = gpu_eval(prog_words, cols, n_rows, trace_n, coeff, denom, acc, entry)
            for c in range(4):
                assert np.array_equal(got[c], want[c]), (jit, c)
    finally:
        call("vx_bytecode_set_jit", 1)


def test_reference_vectors():
    one = [[1, 0, 0, 0]]
    w = [(0x03 << 24) | 0, 0, 0, (0x01 << 24) | 1, 1, (0x11 << 24) | 2, (1 << 16) | 0, (0x12 << 24) | 3, 2,
         (0x40 << 24) | 3]
    out = gpu_eval(w, [np.array([0, 1, 0, 1], np.uint32)], 4, 4, one, [1])
    assert all(not np.any(o) for o in out)
    w = [(0x03 << 24) | 0, 0, 0, (0x14 << 24) | 1, 0, P - 5, (0x40 << 24) | 1]
    out = gpu_eval(w, [np.array([3, 5, 7, 5], np.uint32)], 4, 4, one, [1])
    assert [int(x) for x in out[0]] == [P - 2, 0, 2, 0] and not np.any(out[1])


@pytest.mark.parametrize("seed,log_rows,log_expand,n_ops",
                         [(1, 2, 0, 60), (2, 6, 1, 200), (3, 10, 2, 300), (4, 16, 1, 150), (5, 12, 3, 400)])
def test_random_programs_match_oracle(seed, log_rows, log_expand, n_ops):
    rng = random.Random(seed)
    n_rows, n_cols = 1 << log_rows, 7
    trace_n = n_rows >> log_expand
    offs = (0,) if log_expand == 0 else (0, 0, 1, -1, 3, -2)
    prog = pb.random_program(rng, n_ops, n_cols, n_regs=24, offsets=offs)
    npr = np.random.default_rng(seed)
    cols = [npr.integers(0, P, n_rows, dtype=np.uint32) for _ in range(n_cols)]
    cols[2] = None
    coeff = [pb.rand_q(rng) for _ in range(prog.n_constraints())]
    denom = [rng.randrange(1, P) for _ in range(1 << log_expand)]
    print("bytecode seed", seed, "rows", n_rows, "ops", len(prog.ops), flush=True)
    check(prog.encode(), cols, n_rows, trace_n, coeff, denom)


@pytest.mark.parametrize("live,in_lds", [(6, True), (90, False)])
def test_register_file_in_lds_and_in_workspace(live, in_lds):
    # `live` QM31 values held at once: 6 (32 words per row with the four
    # trace values) fit the LDS register file (<= 128 words), 90 (364 words,
    # trace values re-read at each use) need the workspace
    import ctypes as C
    rng = random.Random(live)
    p = pb.Prog()
    for r in range(live):
        p.load_trace(r, r % 4, 0)
        p.base_mul_secure_const(1000 + r, r, pb.rand_q(rng))
    acc = 2000
    p.widen(acc, 0)
    for r in range(live):
        p.wide_mul(acc, acc, 1000 + r)
        p.add_constraint(1000 + r)
    p.add_constraint(acc)
    words = np.array(p.encode(), np.uint32)
    m, q, k = C.c_uint32(), C.c_uint32(), C.c_uint32()
    call("vx_bytecode_slots", words.ctypes.data, words.size, 4, C.byref(m), C.byref(q), C.byref(k))
    assert (m.value + 4 * q.value <= 128) == in_lds
    n_rows = 1 << 11
    npr = np.random.default_rng(live)
    cols = [npr.integers(0, P, n_rows, dtype=np.uint32) for _ in range(4)]
    coeff = [pb.rand_q(rng) for _ in range(p.n_constraints())]
    print("bytecode live values", live, flush=True)
    check(p.encode(), cols, n_rows, n_rows // 2, coeff, [rng.randrange(1, P), rng.randrange(1, P)])


def test_accumulates_and_warp_entry_matches():
    rng = random.Random(11)
    prog = pb.random_program(rng, 150, 3, offsets=(0, 1, -1))
    n_rows = 1 << 9
    npr = np.random.default_rng(11)
    cols = [npr.integers(0, P, n_rows, dtype=np.uint32) for _ in range(3)]
    coeff = [pb.rand_q(rng) for _ in range(prog.n_constraints())]
    denom = [5, 6, 7, 8]
    acc = [npr.integers(0, P, n_rows, dtype=np.uint32) for _ in range(4)]
    check(prog.encode(), cols, n_rows, n_rows // 4, coeff, denom, acc)
    check(prog.encode(), cols, n_rows, n_rows // 4, coeff, denom, acc, entry="vx_warp_bytecode_constraint_eval")


def test_rejected_program_writes_nothing():
    n_rows = 64
    acc = [np.full(n_rows, 9, np.uint32) for _ in range(4)]
    outs = [dev(a) for a in acc]
    w = dev(np.array([(0x03 << 24) | 0, 4, 0, (0x40 << 24) | 0], np.uint32))  # column 4 of 1
    _, table = col_table([np.zeros(n_rows, np.uint32)])
    k, dn = dev(np.array([1, 0, 0, 0], np.uint32)), dev(np.array([1], np.uint32))
    with pytest.raises(VortexError):
        call("vx_bytecode_constraint_eval", w.ptr, 4, table.ptr, None, 1, n_rows, n_rows, k.ptr, dn.ptr, 0,
             *[o.ptr for o in outs], 64, None)
    sync()
    assert all(np.array_equal(host(o), a) for o, a in zip(outs, acc))
