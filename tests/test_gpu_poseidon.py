"""Poseidon252 operations of the stwo plugin on the GPU (vx_grind_pow_poseidon,
vx_build_leaves_poseidon252, vx_build_next_layer_poseidon252, replacing
cuda_grind_pow_poseidon and build_leaves/next_layer_poseidon252,
src/cuda/ffi.rs:948-974,1205-1212) against the oracle, bit-exact: the
reference's own test layouts (crates/cuda-backend/src/tests.rs:1803-1882),
every sponge remainder case of the column count, uneven lifted columns,
partial leaf ranges, grind batches with and without a hit, and the whole
search. Elements cross the boundary as 4 u64 limbs of their Montgomery
form."""
import ctypes as C
import random

import numpy as np
import pytest

import poseidon252 as ps
from gpu_util import dev, empty, host, ptr, sync
from vortexstark_amd._lib import VortexError, call

pytestmark = pytest.mark.gpu
NULL = None
P = ps.P
M31 = 0x7FFFFFFF


def felt_words(xs):
    """Canonical values -> device words of their Montgomery forms."""
    out = []
    for x in xs:
        m = ps.to_mont(x)
        out += [(m >> (32 * k)) & 0xFFFFFFFF for k in range(8)]
    return np.array(out, np.uint32)


def words_felts(a):
    a = [int(w) for w in a]
    return [ps.from_mont(sum(a[8 * i + k] << (32 * k) for k in range(8))) for i in range(len(a) // 8)]


def col_table(cols):
    """Device columns plus the device pointer and log-size arrays the plugin uploads."""
    bufs = [dev(np.array(c, np.uint32)) for c in cols]
    ptrs = dev(np.array([b.ptr for b in bufs], np.uint64).view(np.uint32)) if cols else None
    logs = dev(np.array([len(c).bit_length() - 1 for c in cols], np.uint32)) if cols else None
    return bufs, ptrs, logs


def run_leaves(cols, lifting, n_leaves=None):
    n_leaves = (1 << lifting) if n_leaves is None else n_leaves
    bufs, dptrs, dlogs = col_table(cols)
    out = empty(8 * n_leaves)
    call("vx_build_leaves_poseidon252", ptr(dptrs) if dptrs else NULL, ptr(dlogs) if dlogs else NULL, len(cols),
         lifting, ptr(out), n_leaves, NULL)
    return words_felts(host(out))


def test_next_layer_reference_vector():
    # poseidon_hash(1, 2) as the reference's test builds it (tests.rs:1851-1882)
    prev = dev(felt_words([1, 2]))
    out = empty(8)
    call("vx_build_next_layer_poseidon252", ptr(prev), ptr(out), 1, NULL)
    assert words_felts(host(out)) == [ps.hash2(1, 2)]


@pytest.mark.parametrize("n_parents", [1, 7, 300, 1000])
def test_next_layer(n_parents):
    print("poseidon next layer", n_parents, flush=True)
    rng = random.Random(n_parents)
    vals = [rng.randrange(P) for _ in range(2 * n_parents - 3)] + [0, 1, P - 1]
    prev, out = dev(felt_words(vals)), empty(8 * n_parents)
    call("vx_build_next_layer_poseidon252", ptr(prev), ptr(out), n_parents, NULL)
    assert words_felts(host(out)) == ps.next_layer(vals)


def test_leaves_reference_layout():
    # tests.rs:1803-1848: 4 columns of 8 rows, value c * 100 + r + 1, lifting 3
    cols = [[c * 100 + r + 1 for r in range(8)] for c in range(4)]
    assert run_leaves(cols, 3) == ps.leaf_hashes(cols, 3)


@pytest.mark.parametrize("n_cols", [0, 1, 7, 8, 9, 16, 17, 24, 33])
def test_leaves_every_remainder(n_cols):
    print("poseidon leaves", n_cols, flush=True)
    rng = random.Random(100 + n_cols)
    lifting = 5
    cols = [[rng.randrange(M31) for _ in range(1 << rng.randint(1, lifting))] for _ in range(n_cols)]
    cols.sort(key=len)
    assert run_leaves(cols, lifting) == ps.leaf_hashes(cols, lifting)


def test_leaves_uneven_large_and_partial_range():
    rng = random.Random(7)
    lifting = 12
    logs = [3, 6, 9, 10, 12, 12, 12, 11, 1, 12]
    cols = [[rng.randrange(M31) for _ in range(1 << L)] for L in logs]
    want = ps.leaf_hashes(cols, lifting)
    assert run_leaves(cols, lifting) == want
    assert run_leaves(cols, lifting, n_leaves=1000) == want[:1000]


def grind_batch(prefixed, bits, offset, n):
    dp = dev(felt_words([prefixed]))
    res = dev(np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32))
    call("vx_grind_pow_poseidon", ptr(dp), ptr(res), bits, offset, n, NULL)
    r = host(res)
    v = int(r[0]) | (int(r[1]) << 32)
    return None if v == (1 << 64) - 1 else v


@pytest.mark.parametrize("bits,offset,n", [(6, 0, 4096), (6, 1000, 3000), (9, 37, 2000), (20, 0, 500)])
def test_grind_batch(bits, offset, n):
    print("poseidon grind batch", bits, offset, n, flush=True)
    rng = random.Random(bits * 1000 + offset)
    prefixed = ps.pow_prefix(rng.randrange(P), bits)
    assert grind_batch(prefixed, bits, offset, n) == ps.grind_min(prefixed, bits, offset, n)


@pytest.mark.parametrize("bits", [1, 8, 11])
def test_grind_whole_search(bits):
    print("poseidon grind search", bits, flush=True)
    rng = random.Random(50 + bits)
    prefixed = ps.pow_prefix(rng.randrange(P), bits)
    dp, res = dev(felt_words([prefixed])), empty(2)
    call("vx_grind_pow_stwo_poseidon", ptr(dp), bits, ptr(res), NULL)
    r = host(res)
    got = int(r[0]) | (int(r[1]) << 32)
    assert got == ps.grind_min(prefixed, bits, 0, 1 << 16)
    assert ps.pow_zeros(prefixed, got) >= bits


def test_grind_zero_bits_and_bad_arguments():
    dp, res = dev(felt_words([5])), dev(np.array([7, 7], np.uint32))
    call("vx_grind_pow_stwo_poseidon", ptr(dp), 0, ptr(res), NULL)
    assert list(host(res)) == [0, 0]
    with pytest.raises(VortexError, match="pow_bits"):
        call("vx_grind_pow_stwo_poseidon", ptr(dp), 41, ptr(res), NULL)
    with pytest.raises(VortexError, match="null"):
        call("vx_build_next_layer_poseidon252", NULL, ptr(res), 1, NULL)
    out = empty(64)
    with pytest.raises(VortexError, match="n_leaves"):
        call("vx_build_leaves_poseidon252", NULL, NULL, 0, 2, ptr(out), 5, NULL)
    sync()
