"""The allocation-free lifted Merkle root (vx_merkle_root_lifted_noalloc)
against the allocating entry point (vx_merkle_root_lifted) through the C-ABI:
two calls in a row share the device scratch, the second with a smaller input,
then a larger one again; small trees also against the oracle. Shapes whose
scratch need exceeds the fixed maximum are rejected on the host."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import vortexstark_amd as vx
from gpu_util import dev, empty, host, sync

pytestmark = pytest.mark.gpu


def ptrs(bufs):
    return (C.c_void_p * len(bufs))(*[b.ptr for b in bufs])


def u32s(xs):
    return (C.c_uint32 * len(xs))(*xs)


def columns(logs, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, po.P, size=1 << L, dtype=np.uint32) for L in logs]


def both_roots(logs, seed):
    cols = columns(logs, seed)
    bufs = [dev(c) for c in cols]
    want, got = empty(8), empty(8)
    vx._lib.call("vx_merkle_root_lifted", ptrs(bufs), u32s(logs), len(logs), want.ptr, None)
    vx._lib.call("vx_merkle_root_lifted_noalloc", ptrs(bufs), u32s(logs), len(logs), got.ptr, None)
    sync()
    return cols, host(want), host(got)


def test_noalloc_lifted_root_matches_allocating_entry_point():
    # first call: 40 columns up to 2^18 rows (1024-leaf tiles, 4 leaves per lane)
    big = sorted([14, 18, 12, 16, 9, 18, 17, 15] * 5)
    _, want, got = both_roots(big, 1)
    print("noalloc call 1: cols", len(big), "log", big[-1], flush=True)
    assert np.array_equal(got, want)
    # second call, smaller: 5 columns up to 2^12 rows (256-leaf tiles), scratch
    # still holds the first call's table and levels
    small = [4, 6, 9, 12, 12]
    cols, want, got = both_roots(small, 2)
    print("noalloc call 2: cols", len(small), "log", small[-1], flush=True)
    assert np.array_equal(got, want)
    assert np.array_equal(got, po.merkle_root_lifted(cols))
    # third call, larger again, with more columns than one table chunk
    many = sorted([5 + (i % 16) for i in range(150)])
    _, want, got = both_roots(many, 3)
    print("noalloc call 3: cols", len(many), "log", many[-1], flush=True)
    assert np.array_equal(got, want)


def test_noalloc_lifted_root_largest_accepted_tree():
    # one 2^26-row column: 4 MiB of stored levels, within the 8 MiB scratch
    logs = [20, 26]
    _, want, got = both_roots(logs, 4)
    assert np.array_equal(got, want)


def test_noalloc_lifted_root_rejects_oversize():
    # 2^27 rows need 8 MiB of levels plus the table: rejected before any
    # device work (the column pointers are never read)
    fake = (C.c_void_p * 1)(0x1000)
    out = empty(8)
    with pytest.raises(vx._lib.VortexError, match="8 MiB"):
        vx._lib.call("vx_merkle_root_lifted_noalloc", fake, u32s([27]), 1, out.ptr, None)
    with pytest.raises(vx._lib.VortexError, match="4096"):
        cols = (C.c_void_p * 4097)(*([0x1000] * 4097))
        vx._lib.call("vx_merkle_root_lifted_noalloc", cols, u32s([4] * 4097), 4097, out.ptr, None)
