"""Independent pure-Python restatement of the Fibonacci Circle-STARK prove path.

Written from the reference source text (cited per function), separately from
the C oracle, so that the two can be checked against each other. Hashing uses
Python's hashlib.blake2s (standard RFC 7693 Blake2s-256), field arithmetic
uses Python integers, and twiddles are derived from closed-form point powers
rather than by the iterated squashing the C oracle uses. Only small sizes are
practical (log_n <= 8).

Used only by tests/golden/gen_golden.py to write fixtures and by CPU tests.
"""
import hashlib
import struct

P = (1 << 31) - 1
GEN = (2, 1268011823)  # src/circle.rs:26-29
IDENTITY = (1, 0)


# ---------------- field: src/field/{m31,cm31,qm31}.rs ----------------

def inv(a):
    return pow(a, P - 2, P)  # m31.rs:32-35 (inv(0) = 0)


def cm_mul(x, y):
    (a, b), (c, d) = x, y
    return ((a * c - b * d) % P, (a * d + b * c) % P)


def qm_add(x, y):
    return tuple((u + v) % P for u, v in zip(x, y))


def qm_sub(x, y):
    return tuple((u - v) % P for u, v in zip(x, y))


def qm_mul(x, y):
    # qm31.rs:113-132 with u^2 = 2 + i
    a, b = (x[0], x[1]), (x[2], x[3])
    c, d = (y[0], y[1]), (y[2], y[3])
    ac, bd = cm_mul(a, c), cm_mul(b, d)
    ad, bc = cm_mul(a, d), cm_mul(b, c)
    bd_u2 = cm_mul(bd, (2, 1))
    return ((ac[0] + bd_u2[0]) % P, (ac[1] + bd_u2[1]) % P,
            (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P)


def qm_scale(x, s):
    return tuple(v * s % P for v in x)


# ---------------- circle: src/circle.rs:12-185 ----------------

def pt_mul(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def pt_pow(p, k):
    r = IDENTITY
    while k:
        if k & 1:
            r = pt_mul(r, p)
        p = pt_mul(p, p)
        k >>= 1
    return r


def gen_pow2(e):
    """G^(2^e)"""
    return pt_pow(GEN, 1 << e)


def subgroup(L):
    return (IDENTITY, gen_pow2(31 - L), L)


def half_coset(L):
    return (gen_pow2(30 - L), gen_pow2(31 - L), L)


def half_odds(L):
    return (gen_pow2(max(29 - L, 0)), gen_pow2(31 - L), L)


def eval_domain(L):
    return half_coset(L) if L <= 30 else subgroup(L)


def coset_at(c, i):
    return pt_mul(c[0], pt_pow(c[1], i))


# ---------------- twiddles (closed form) ----------------
# circle.rs:284-347 builds T_l[h] by iterated x -> 2x^2-1 over even entries.
# The doubling map sends x(p) to x(p^2), so T_l[h] = x((p_{2^(l+1) h})^(2^l)).

def twiddle_tables(c):
    init, step, L = c
    n = 1 << L
    circle = [coset_at(c, i)[1] for i in range(n // 2)]
    line = []
    for l in range(L - 1):
        size = n >> (l + 1)
        line.append([pt_pow(coset_at(c, (1 << (l + 1)) * h), 1 << l)[0] for h in range(size)])
    return line, circle


# ---------------- NTT: cuda/circle_ntt.cu:74-117, 248-312 ----------------

def _layer(v, tw, j, forward):
    s = 1 << j
    out = list(v)
    for start in range(0, len(v), 2 * s):
        w = tw[start // (2 * s)]
        for k in range(s):
            a, b = out[start + k], out[start + k + s]
            if forward:
                t = b * w % P
                out[start + k], out[start + k + s] = (a + t) % P, (a - t) % P
            else:
                out[start + k], out[start + k + s] = (a + b) % P, (a - b) * w % P
    return out


def evaluate(v, c):
    line, circle = twiddle_tables(c)
    L = c[2]
    for l in reversed(range(L - 1)):
        v = _layer(v, line[l], l + 1, True)
    return _layer(v, circle, 0, True)


def interpolate(v, c):
    line, circle = twiddle_tables(c)
    L = c[2]
    v = _layer(v, [inv(t) for t in circle], 0, False)
    for l in range(L - 1):
        v = _layer(v, [inv(t) for t in line[l]], l + 1, False)
    scale = inv(1 << L)  # equals 2^((30 L) mod 31)
    return [x * scale % P for x in v]


# ---------------- hashing / Merkle: cuda/blake2s.cu:13-87 ----------------

def le_words(ws):
    return b"".join(struct.pack("<I", w) for w in ws)


def words(b):
    return list(struct.unpack("<8I", b))


def b2(data):
    return hashlib.blake2s(data).digest()


def merkle_layers(leaves_words):
    layer = [b2(le_words(w)) for w in leaves_words]
    layers = [layer]
    while len(layer) > 1:
        layer = [b2(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
        layers.append(layer)
    return layers


def merkle_layers_from_hashes(leaf_hashes):
    """Tree over precomputed 32-byte leaf hashes (node = Blake2s(left || right))."""
    layer = list(leaf_hashes)
    layers = [layer]
    while len(layer) > 1:
        layer = [b2(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
        layers.append(layer)
    return layers


def auth_path(layers, idx):
    path = []
    for layer in layers[:-1]:
        path.append(words(layer[idx ^ 1]))
        idx //= 2
    return path


# ---------------- channel: src/channel.rs:15-185 ----------------

class Channel:
    def __init__(self):
        self.state = bytes(32)
        self.n_draws = 0

    def mix_digest(self, d):
        self.state = b2(self.state + le_words(d))
        self.n_draws = 0

    def mix_felts(self, felts):
        self.state = b2(self.state + b"".join(le_words(f) for f in felts))
        self.n_draws = 0

    def mix_u64(self, v):
        self.state = b2(self.state + struct.pack("<Q", v))
        self.n_draws = 0

    def squeeze(self):
        out = b2(self.state + struct.pack("<I", self.n_draws) + b"\x00")
        self.n_draws += 1
        return out

    def draw_felt(self):
        while True:
            w = struct.unpack("<8I", self.squeeze())
            if all(x < 2 * P for x in w):
                return tuple(x % P for x in w[:4])

    def draw_number(self, bound):
        return struct.unpack("<Q", self.squeeze()[:8])[0] % bound

    def pow_ok(self, bits, nonce):
        prefix = b2(struct.pack("<I", 0x12345678) + bytes(12) + self.state + struct.pack("<I", bits))
        h = b2(prefix + struct.pack("<Q", nonce))
        v = int.from_bytes(h[:16], "little")
        tz = 128 if v == 0 else (v & -v).bit_length() - 1
        return tz >= bits


# ---------------- prover: src/prover.rs:264-619 ----------------

def fib_trace(a, b, log_n):
    t = [a, b]
    while len(t) < (1 << log_n):
        t.append((t[-1] + t[-2]) % P)
    return t


def bitrev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def fold_tw(c, i, use_y):
    p = coset_at(c, bitrev(2 * i, c[2]))
    return inv(p[1] if use_y else p[0])


def fold(f0, f1, alpha, tw):
    return qm_add(qm_add(f0, f1), qm_mul(alpha, qm_scale(qm_sub(f0, f1), tw)))


def prove(a, b, log_n, blowup_bits=2, n_queries=80, pow_bits=0):
    le = log_n + blowup_bits
    N = 1 << le
    coeffs = interpolate(fib_trace(a, b, log_n), half_coset(log_n))
    ev = evaluate(coeffs + [0] * (N - len(coeffs)), eval_domain(le))
    t_layers = merkle_layers([[x] for x in ev])
    ch = Channel()
    ch.mix_digest(words(t_layers[-1][0]))
    alpha = ch.draw_felt()
    q = [qm_scale(alpha, (ev[(i + 2) % N] - ev[(i + 1) % N] - ev[i]) % P) for i in range(N)]
    q_layers = merkle_layers(q)
    ch.mix_digest(words(q_layers[-1][0]))
    alpha0 = ch.draw_felt()
    dom = eval_domain(le)
    layer = [fold(q[2 * i], q[2 * i + 1], alpha0, fold_tw(dom, i, True)) for i in range(N // 2)]
    fri = [(layer, merkle_layers(layer))]
    ch.mix_digest(words(fri[-1][1][-1][0]))
    cur = le - 1
    while cur > 3:
        al = ch.draw_felt()
        d = half_odds(cur)
        layer = [fold(layer[2 * i], layer[2 * i + 1], al, fold_tw(d, i, False))
                 for i in range(len(layer) // 2)]
        cur -= 1
        fri.append((layer, merkle_layers(layer)))
        ch.mix_digest(words(fri[-1][1][-1][0]))
    ch.mix_felts(layer)
    nonce = 0
    if pow_bits:
        while not ch.pow_ok(pow_bits, nonce):
            nonce += 1
        ch.mix_u64(nonce)
    qis = [ch.draw_number(N) for _ in range(n_queries)]

    out = [0x4B52414B, 2 if pow_bits else 1, log_n, a, b]
    if pow_bits:
        out += [blowup_bits, pow_bits, nonce & 0xFFFFFFFF, nonce >> 32]
    out += words(t_layers[-1][0]) + words(q_layers[-1][0])
    out += [len(fri)] + sum((words(f[1][-1][0]) for f in fri), [])
    out += [len(layer)] + sum((list(v) for v in layer), [])
    out += [len(qis)] + qis

    def decom(vals, layers, idxs):
        res = [len(idxs)]
        for i in idxs:
            res += list(vals[i])
        for i in idxs:
            res += list(vals[i ^ 1])
        for sel in (0, 1):
            for i in idxs:
                pth = auth_path(layers, i ^ sel)
                res += [len(pth)] + sum(pth, [])
        return res

    out += decom([[x] for x in ev], t_layers, qis)
    out += decom(q, q_layers, qis)
    out += [len(fri)]
    idx = [i // 2 for i in qis]
    for vals, layers in fri:
        out += decom(vals, layers, idx)
        idx = [i // 2 for i in idx]
    return le_words(out)


# ---------------- synthetic degree-2 multi-column AIR (build-defined) ----------------
# Restated from the definition (not from oracle/vx_oracle.c): W columns of
# SplitMix64-derived values, K constraints C_k = c_k t_a t_b' + d_k t_e with
# a = k mod W, b = (5k + 1) mod W, e = (11k + 7) mod W; Q = sum alpha^(k+1) C_k;
# trace committed in trees of <= 16 columns; KRAK v3 bytes.

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def d2_value(seed, col, row):
    return mix64((seed + GOLDEN * (((col << 32) | row) + 1)) & M64) % P


def d2_coeffs(air_seed, k_cons):
    s, out = air_seed, []
    for _ in range(2 * k_cons):
        s = (s + GOLDEN) & M64
        out.append(mix64(s) % P)
    return out


def d2_q(row, nxt, w, k_cons, cd, alpha):
    acc, beta = (0, 0, 0, 0), alpha
    for k in range(k_cons):
        m = (cd[2 * k] * row[k % w] * nxt[(5 * k + 1) % w] + cd[2 * k + 1] * row[(11 * k + 7) % w]) % P
        acc = qm_add(acc, qm_scale(beta, m))
        beta = qm_mul(beta, alpha)
    return acc


def prove_d2(seed, log_n, w=34, k_cons=35, air_seed=0, blowup_bits=2, n_queries=80, pow_bits=0):
    le = log_n + blowup_bits
    N = 1 << le
    hc, ed = half_coset(log_n), eval_domain(le)
    cols = []
    for c in range(w):
        coeffs = interpolate([d2_value(seed, c, i) for i in range(1 << log_n)], hc)
        cols.append(evaluate(coeffs + [0] * (N - len(coeffs)), ed))
    groups = [list(range(g, min(g + 16, w))) for g in range(0, w, 16)]
    t_layers = [merkle_layers([[cols[c][i] for c in grp] for i in range(N)]) for grp in groups]
    ch = Channel()
    for tl in t_layers:
        ch.mix_digest(words(tl[-1][0]))
    alpha = ch.draw_felt()
    cd = d2_coeffs(air_seed, k_cons)
    q = [d2_q([cols[c][i] for c in range(w)], [cols[c][(i + 1) % N] for c in range(w)], w, k_cons, cd, alpha)
         for i in range(N)]
    q_layers = merkle_layers(q)
    ch.mix_digest(words(q_layers[-1][0]))
    alpha0 = ch.draw_felt()
    layer = [fold(q[2 * i], q[2 * i + 1], alpha0, fold_tw(ed, i, True)) for i in range(N // 2)]
    fri = [(layer, merkle_layers(layer))]
    ch.mix_digest(words(fri[-1][1][-1][0]))
    cur = le - 1
    while cur > 3:
        al = ch.draw_felt()
        d = half_odds(cur)
        layer = [fold(layer[2 * i], layer[2 * i + 1], al, fold_tw(d, i, False)) for i in range(len(layer) // 2)]
        cur -= 1
        fri.append((layer, merkle_layers(layer)))
        ch.mix_digest(words(fri[-1][1][-1][0]))
    ch.mix_felts(layer)
    nonce = 0
    if pow_bits:
        while not ch.pow_ok(pow_bits, nonce):
            nonce += 1
        ch.mix_u64(nonce)
    qis = [ch.draw_number(N) for _ in range(n_queries)]

    out = [0x4B52414B, 3, log_n, w, k_cons, blowup_bits, pow_bits, nonce & 0xFFFFFFFF, nonce >> 32,
           seed & 0xFFFFFFFF, seed >> 32, air_seed & 0xFFFFFFFF, air_seed >> 32, len(groups)]
    for tl in t_layers:
        out += words(tl[-1][0])
    out += words(q_layers[-1][0])
    out += [len(fri)] + sum((words(f[1][-1][0]) for f in fri), [])
    out += [len(layer)] + sum((list(v) for v in layer), [])
    out += [len(qis)] + qis
    for grp, tl in zip(groups, t_layers):
        nx = [(i + 1) % N for i in qis]
        out += [len(qis)]
        for idxs in (qis, nx):
            for i in idxs:
                out += [cols[c][i] for c in grp]
        for idxs in (qis, nx):
            for i in idxs:
                pth = auth_path(tl, i)
                out += [len(pth)] + sum(pth, [])

    def decom(vals, layers, idxs):
        res = [len(idxs)]
        for i in idxs:
            res += list(vals[i])
        for i in idxs:
            res += list(vals[i ^ 1])
        for sel in (0, 1):
            for i in idxs:
                pth = auth_path(layers, i ^ sel)
                res += [len(pth)] + sum(pth, [])
        return res

    out += decom(q, q_layers, qis)
    out += [len(fri)]
    idx = [i // 2 for i in qis]
    for vals, layers in fri:
        out += decom(vals, layers, idx)
        idx = [i // 2 for i in idx]
    return le_words(out)


# ---------------- batched-columns proof (KRAK v4, build-defined) ----------------
# Columns c of log sizes L_c (ascending), values d2_value(seed, c, i), LDE at
# blowup, bit reversed; lifted tree (leaf = Blake2s of every column at its
# lifted row); the channel mixes Blake2s(log sizes) then the root; Q(i) =
# sum_c alpha^(c+1) col_c[lifted row]; then the FRI phase.

def lifted_row(i, S, L):
    return i if L == S else ((i >> (S - L + 1)) << 1) | (i & 1)


def prove_batch(seed, log_sizes, blowup_bits=2, n_queries=80, pow_bits=0):
    elog = [L + blowup_bits for L in log_sizes]
    S = elog[-1]
    N = 1 << S
    cols = []
    for c, L in enumerate(log_sizes):
        coeffs = interpolate([d2_value(seed, c, i) for i in range(1 << L)], half_coset(L))
        ev = evaluate(coeffs + [0] * ((1 << elog[c]) - len(coeffs)), eval_domain(elog[c]))
        cols.append([ev[bitrev(i, elog[c])] for i in range(len(ev))])

    def leaf_words(i):
        return [cols[c][lifted_row(i, S, elog[c])] for c in range(len(cols))]

    t_layers = merkle_layers_from_hashes([b2(le_words(leaf_words(i))) for i in range(N)])
    ch = Channel()
    ch.mix_digest(words(b2(le_words(list(log_sizes)))))   # the sizes enter the transcript first
    ch.mix_digest(words(t_layers[-1][0]))
    alpha = ch.draw_felt()
    betas, b = [], alpha
    for _ in cols:
        betas.append(b)
        b = qm_mul(b, alpha)
    q = []
    for i in range(N):
        acc = (0, 0, 0, 0)
        for be, v in zip(betas, leaf_words(i)):
            acc = qm_add(acc, qm_scale(be, v))
        q.append(acc)
    q_layers = merkle_layers(q)
    ch.mix_digest(words(q_layers[-1][0]))
    alpha0 = ch.draw_felt()
    ed = eval_domain(S)
    layer = [fold(q[2 * i], q[2 * i + 1], alpha0, fold_tw(ed, i, True)) for i in range(N // 2)]
    fri = [(layer, merkle_layers(layer))]
    ch.mix_digest(words(fri[-1][1][-1][0]))
    cur = S - 1
    while cur > 3:
        al = ch.draw_felt()
        d = half_odds(cur)
        layer = [fold(layer[2 * i], layer[2 * i + 1], al, fold_tw(d, i, False)) for i in range(len(layer) // 2)]
        cur -= 1
        fri.append((layer, merkle_layers(layer)))
        ch.mix_digest(words(fri[-1][1][-1][0]))
    ch.mix_felts(layer)
    nonce = 0
    if pow_bits:
        while not ch.pow_ok(pow_bits, nonce):
            nonce += 1
        ch.mix_u64(nonce)
    qis = [ch.draw_number(N) for _ in range(n_queries)]

    out = [0x4B52414B, 4, len(cols), blowup_bits, pow_bits, nonce & 0xFFFFFFFF, nonce >> 32,
           seed & 0xFFFFFFFF, seed >> 32] + list(log_sizes)
    out += words(t_layers[-1][0]) + words(q_layers[-1][0])
    out += [len(fri)] + sum((words(f[1][-1][0]) for f in fri), [])
    out += [len(layer)] + sum((list(v) for v in layer), [])
    out += [len(qis)] + qis
    out += [len(qis)]
    for side in (0, 1):
        for i in qis:
            out += leaf_words(i ^ side)
    for side in (0, 1):
        for i in qis:
            pth = auth_path(t_layers, i ^ side)
            out += [len(pth)] + sum(pth, [])

    def decom(vals, layers, idxs):
        res = [len(idxs)]
        for i in idxs:
            res += list(vals[i])
        for i in idxs:
            res += list(vals[i ^ 1])
        for sel in (0, 1):
            for i in idxs:
                pth = auth_path(layers, i ^ sel)
                res += [len(pth)] + sum(pth, [])
        return res

    out += decom(q, q_layers, qis)
    out += [len(fri)]
    idx = [i // 2 for i in qis]
    for vals, layers in fri:
        out += decom(vals, layers, idx)
        idx = [i // 2 for i in idx]
    return le_words(out)
