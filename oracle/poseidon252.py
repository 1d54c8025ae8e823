"""Poseidon252 over the Stark field, restated in pure Python (big integers).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker for the device
Poseidon252 grind and Merkle kernels. The product never imports it.

Follows the reference's device code for the stwo plugin's Poseidon252
operations:
- the permutation (Starknet Hades: width 3, 4 full + 83 partial + 4 full
  rounds, x^3, MDS [[3,1,1],[1,-1,1],[1,1,-2]]) of
  cuda/include/poseidon252.cuh:35-82, here in its plain uncompressed form;
  the round constants are sha256("Hades" + j) mod p, the published rule of
  the parameters the reference takes from starknet-crypto-codegen 0.3.3
  (cuda/dump_poseidon_comp_consts/src/main.rs:1-63);
- poseidon_hash(x, y) = permute([x, y, 2])[0] (poseidon252.cuh:83-89);
- the sponge poseidon_hash_many (cuda/merkle_poseidon252.cu:3-11);
- M31 packing into a field element (poseidon252.cuh:94-120);
- lifted leaf hashes and parent layers (merkle_poseidon252.cu:28-148);
- the grind test (cuda/grind_poseidon.cu:1-21,36-75).

Pinned by tests/test_oracle_poseidon.py: the reference's known answer for
the permutation (main.rs:57-62), the Montgomery forms of 1 and 2 that the
reference's header states (poseidon252.cuh:20-28), and closed forms for the
packing, the sponge's padding cases and the grind's trailing-zero count.
"""
import hashlib

P = 2**251 + 17 * 2**192 + 1
R = 2**256 % P           # Montgomery factor, 2^256 mod p
RINV = pow(2**256, -1, P)
FULL_HALF, PARTIAL = 4, 83
ROUNDS = 2 * FULL_HALF + PARTIAL
ARK = [[int(hashlib.sha256(("Hades%d" % (3 * r + j)).encode()).hexdigest(), 16) % P for j in range(3)]
       for r in range(ROUNDS)]


def mix(s):
    t = (s[0] + s[1] + s[2]) % P
    return [(t + 2 * s[0]) % P, (t - 2 * s[1]) % P, (t - 3 * s[2]) % P]


def permute(s):
    s = [x % P for x in s]
    for r in range(ROUNDS):
        s = [(s[j] + ARK[r][j]) % P for j in range(3)]
        if FULL_HALF <= r < FULL_HALF + PARTIAL:
            s[2] = pow(s[2], 3, P)
        else:
            s = [pow(x, 3, P) for x in s]
        s = mix(s)
    return s


def hash2(x, y):
    """poseidon_hash(x, y)."""
    return permute([x, y, 2])[0]


def hash_many(msgs):
    """poseidon_hash_many: absorb pairs; pad a lone element with 1 in the
    second rate word, an even count with 1 in the first."""
    s = [0, 0, 0]
    n2 = len(msgs) // 2 * 2
    for k in range(0, n2, 2):
        s[0] = (s[0] + msgs[k]) % P
        s[1] = (s[1] + msgs[k + 1]) % P
        s = permute(s)
    if len(msgs) % 2:
        s[0] = (s[0] + msgs[-1]) % P
        s[1] = (s[1] + 1) % P
    else:
        s[0] = (s[0] + 1) % P
    return permute(s)[0]


def pack_m31s(vals):
    """Up to 8 M31 values, first one most significant, 31 bits each; fewer
    than 8 add their count times 2^248."""
    assert 1 <= len(vals) <= 8
    x = 0
    for v in vals:
        x = (x << 31) | int(v)
    if len(vals) < 8:
        x += len(vals) << 248
    return x


def lifted_row(leaf, lifting_log, col_log):
    if col_log == lifting_log:
        return leaf
    shift = lifting_log - col_log + 1
    return ((leaf >> shift) << 1) | (leaf & 1)


def leaf_hashes(cols, lifting_log, n_leaves=None):
    """Leaf i: hash_many of the packed values of every column at its lifted
    row, in column order, 8 values per element."""
    n_leaves = (1 << lifting_log) if n_leaves is None else n_leaves
    out = []
    for leaf in range(n_leaves):
        vals = [int(c[lifted_row(leaf, lifting_log, len(c).bit_length() - 1)]) for c in cols]
        felts = [pack_m31s(vals[k:k + 8]) for k in range(0, len(vals), 8)]
        out.append(hash_many(felts))
    return out


def next_layer(prev):
    return [hash2(prev[2 * i], prev[2 * i + 1]) for i in range(len(prev) // 2)]


def pow_zeros(prefixed, nonce):
    """Trailing zeros of the low 128 bits of poseidon_hash(prefixed, nonce)."""
    low = hash2(prefixed, nonce) & ((1 << 128) - 1)
    return 128 if low == 0 else (low & -low).bit_length() - 1


def pow_prefix(digest, pow_bits, pow_prefix_word=0x12345678):
    """poseidon_hash_many([POW_PREFIX, digest, n_bits]) (grind_ops.rs:273-277)."""
    return hash_many([pow_prefix_word, digest, pow_bits])


def grind_min(prefixed, pow_bits, start=0, count=1 << 20):
    """Smallest nonce in [start, start + count) with pow_zeros >= pow_bits, or None."""
    for n in range(start, start + count):
        if pow_zeros(prefixed, n) >= pow_bits:
            return n
    return None


# ------------------------------------------------------------ representations
def to_mont(x):
    return x * 2**256 % P


def from_mont(x):
    return x * RINV % P


def to_limbs(x):
    """Four little-endian 64-bit limbs (the device layout of one element)."""
    return [(x >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(4)]


def from_limbs(limbs):
    return sum(int(v) << (64 * k) for k, v in enumerate(limbs))
