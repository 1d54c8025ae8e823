"""The oracle's Poseidon252 (oracle/poseidon252.py) and the device's round
constant table (vortexstark_amd/csrc/poseidon_consts.h), pinned by:

- the reference's known answer for poseidon_hash
  (cuda/dump_poseidon_comp_consts/src/main.rs:57-62), which fixes the
  round constants, the round structure, the S-box and the MDS;
- the Montgomery forms of 1 and 2 and the first round constant that the
  reference's headers state (cuda/include/poseidon252.cuh:20-28,
  poseidon_round_keys.cuh:15), the closed forms 2^256 mod p, 2^257 mod p;
- closed forms for the M31 packing (8 values fill 248 bits, a short group
  carries its count at bit 248, so groups of different lengths never
  collide) and the grind's trailing-zero count;
- the sponge's padding (merkle_poseidon252.cu:3-11) has no literal vector in
  the reference: parity unpinned beyond the permutation it composes.
"""
import os
import random
import re

import pytest

import poseidon252 as ps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = ps.P


def test_known_answer_poseidon_hash():
    x = 0x0b662f9017fa7956fd70e26129b1833e10ad000fd37b4d9f4e0ce6884b7bbe
    y = 0x01fe356bf76102cdae1bfbdc173602ead228b12904c00dad9cf16e035468bea
    assert ps.hash2(x, y) == 0x75540825a6ecc5dc7d7c2f5f868164182742227f1367d66c43ee51ec7937a81


def test_a_plausible_mistake_fails_the_known_answer(monkeypatch):
    # one partial round fewer, or the MDS's -2 taken as +2, changes the hash
    x, y = 3, 4
    good = ps.hash2(x, y)
    monkeypatch.setattr(ps, "PARTIAL", 82)
    assert ps.hash2(x, y) != good
    monkeypatch.undo()
    monkeypatch.setattr(ps, "mix", lambda s: [(sum(s) + 2 * s[0]) % P, (sum(s) - 2 * s[1]) % P,
                                              (sum(s) + 2 * s[2]) % P])
    assert ps.hash2(x, y) != good


def test_montgomery_forms_match_the_reference_header():
    assert ps.to_limbs(ps.to_mont(1)) == [0xffffffffffffffe1, 0xffffffffffffffff, 0xffffffffffffffff,
                                          0x07fffffffffffdf0]
    assert ps.to_limbs(ps.to_mont(2)) == [0xffffffffffffffc1, 0xffffffffffffffff, 0xffffffffffffffff,
                                          0x07fffffffffffbd0]
    assert ps.to_limbs(ps.to_mont(ps.ARK[0][0])) == [0x8048052ec221e1ea, 0xfb041a7cb2aca4da, 0x62818313eae79571,
                                                     0x04056cdd9992bdde]
    for x in (0, 1, P - 1, 12345):
        assert ps.from_mont(ps.to_mont(x)) == x


def header_words(text, name):
    m = re.search(r"#define %s \{([^}]*)\}" % name, text)
    return [int(w.rstrip("u"), 16) for w in m.group(1).replace(" ", "").split(",")]


def test_device_constant_table_is_the_oracles():
    with open(os.path.join(ROOT, "vortexstark_amd", "csrc", "poseidon_consts.h")) as f:
        text = f.read()

    def val(ws):
        return sum(w << (32 * k) for k, w in enumerate(ws))
    assert val(header_words(text, "VX_F252_MONT_ONE")) == ps.to_mont(1)
    assert val(header_words(text, "VX_F252_MONT_TWO")) == ps.to_mont(2)
    assert val(header_words(text, "VX_F252_R2")) == (2**512) % P
    body = text[text.index("VX_POSEIDON_ARK_MONT"):]
    rows = re.findall(r"\{(0x[^}]*)\}", body)
    assert len(rows) == 3 * ps.ROUNDS
    flat = [c for r in ps.ARK for c in r]
    for j, row in enumerate(rows):
        assert val([int(w.strip().rstrip("u"), 16) for w in row.split(",")]) == ps.to_mont(flat[j]), j


def test_packing_closed_forms():
    rng = random.Random(3)
    v = [rng.randrange(2**31 - 1) for _ in range(8)]
    assert ps.pack_m31s(v) == sum(x << (31 * (7 - i)) for i, x in enumerate(v))
    assert ps.pack_m31s([2**31 - 2] * 8) < 2**248 < P
    for n in range(1, 8):
        assert ps.pack_m31s(v[:n]) == sum(x << (31 * (n - 1 - i)) for i, x in enumerate(v[:n])) + (n << 248)
        assert ps.pack_m31s(v[:n]) < P
    # zero groups of every length are distinct elements
    assert len({ps.pack_m31s([0] * n) for n in range(1, 9)}) == 8


def test_lifted_rows():
    # a column of the tree's size is read at the leaf; a smaller one at the
    # pair (2 * (leaf >> (S - L + 1)) + (leaf & 1)), covering every row
    S = 5
    for L in range(1, S + 1):
        rows = [ps.lifted_row(i, S, L) for i in range(1 << S)]
        assert set(rows) == set(range(1 << L))
        if L == S:
            assert rows == list(range(1 << S))
        for i in range(1 << S):
            assert rows[i] & 1 == i & 1


def test_leaf_hash_is_the_sponge_of_packed_rows():
    cols = [[c * 100 + r + 1 for r in range(8)] for c in range(4)]  # the reference's test layout
    got = ps.leaf_hashes(cols, 3)
    for i in range(8):
        assert got[i] == ps.hash_many([ps.pack_m31s([c[i] for c in cols])])
    # 17 columns: two full groups and one of length 1
    rng = random.Random(5)
    cols = [[rng.randrange(2**31 - 1) for _ in range(4)] for _ in range(17)]
    got = ps.leaf_hashes(cols, 2)
    for i in range(4):
        vals = [c[i] for c in cols]
        assert got[i] == ps.hash_many([ps.pack_m31s(vals[:8]), ps.pack_m31s(vals[8:16]), ps.pack_m31s(vals[16:])])


def test_sponge_cases_compose_the_permutation():
    a, b, c = 11, 22, 33
    assert ps.hash_many([]) == ps.permute([1, 0, 0])[0]
    assert ps.hash_many([a]) == ps.permute([a, 1, 0])[0]
    s = ps.permute([a, b, 0])
    assert ps.hash_many([a, b]) == ps.permute([(s[0] + 1) % P, s[1], s[2]])[0]
    assert ps.hash_many([a, b, c]) == ps.permute([(s[0] + c) % P, (s[1] + 1) % P, s[2]])[0]
    assert ps.hash2(a, b) == ps.permute([a, b, 2])[0]


def test_pow_zeros_and_grind():
    rng = random.Random(9)
    prefixed = ps.pow_prefix(rng.randrange(P), 6)
    for n in range(40):
        low = ps.hash2(prefixed, n) & ((1 << 128) - 1)
        z = 0
        while z < 128 and not (low >> z) & 1:
            z += 1
        assert ps.pow_zeros(prefixed, n) == z
    r = ps.grind_min(prefixed, 6)
    assert r is not None and ps.pow_zeros(prefixed, r) >= 6
    assert all(ps.pow_zeros(prefixed, n) < 6 for n in range(r))
    assert ps.grind_min(prefixed, 6, start=r + 1, count=1) in (None, r + 1)


@pytest.mark.parametrize("n", [2, 6])
def test_next_layer_pairs(n):
    rng = random.Random(n)
    prev = [rng.randrange(P) for _ in range(n)]
    assert ps.next_layer(prev) == [ps.hash2(prev[2 * i], prev[2 * i + 1]) for i in range(n // 2)]
