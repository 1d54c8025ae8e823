"""Programs in the stwo backend's register bytecode and their meaning,
evaluated directly, for the tests of the bytecode constraint evaluator.

The word encoding follows crates/cuda-backend/src/constraint_eval/
bytecode.rs:173-323 (header [opcode:8 | 0:8 | dst:16], 3-register ops
[src1:16 | src2:16], 2-register ops [src:16 | 0:16], constants after the
source word, LOAD_TRACE [flat_col] then [sign:1 | |offset|:31]). `run`
evaluates the op list itself (no decoding) with the field arithmetic of
pyref, and finds a row offset by walking the circle: row i of a bit-reversed
evaluation of 2^e rows is the point at natural index bitrev(i) of stwo's
CircleDomain (half_odds(e - 1), then its conjugates), and `offset` trace
steps multiply it by the trace-domain generator to that power."""
import random

import pyref

P = pyref.P
OPC = {
    "load_const": 0x01, "load_secure_const": 0x02, "load_trace": 0x03,
    "add": 0x10, "sub": 0x11, "mul": 0x12, "neg": 0x13, "add_const": 0x14, "mul_const": 0x15,
    "wide_add": 0x20, "wide_sub": 0x21, "wide_mul": 0x22, "wide_neg": 0x23, "wide_add_const": 0x24,
    "wide_mul_const": 0x25, "wide_add_base": 0x28, "wide_mul_base": 0x29, "base_add_secure_const": 0x2A,
    "base_mul_secure_const": 0x2B, "widen": 0x30, "combine_ef": 0x31, "add_constraint": 0x40,
}
M31_RESULT = {"load_const", "load_trace", "add", "sub", "mul", "neg", "add_const", "mul_const"}


def hdr(op, dst):
    return (OPC[op] << 24) | dst


class Prog:
    def __init__(self):
        self.ops = []

    def __getattr__(self, name):
        if name not in OPC:
            raise AttributeError(name)
        return lambda *args: self.ops.append((name,) + tuple(args))

    def encode(self):
        w = []
        for op, *a in self.ops:
            if op == "load_const":
                w += [hdr(op, a[0]), a[1]]
            elif op == "load_secure_const":
                w += [hdr(op, a[0])] + list(a[1])
            elif op == "load_trace":
                dst, col, off = a
                w += [hdr(op, dst), col, ((1 << 31) | -off) if off < 0 else off]
            elif op in ("add", "sub", "mul", "wide_add", "wide_sub", "wide_mul", "wide_add_base", "wide_mul_base"):
                w += [hdr(op, a[0]), (a[1] << 16) | a[2]]
            elif op in ("neg", "wide_neg", "widen"):
                w += [hdr(op, a[0]), a[1] << 16]
            elif op in ("add_const", "mul_const"):
                w += [hdr(op, a[0]), a[1] << 16, a[2]]
            elif op in ("wide_add_const", "wide_mul_const", "base_add_secure_const", "base_mul_secure_const"):
                w += [hdr(op, a[0]), a[1] << 16] + list(a[2])
            elif op == "combine_ef":
                w += [hdr(op, a[0]), (a[1] << 16) | a[2], (a[3] << 16) | a[4]]
            elif op == "add_constraint":
                w += [hdr(op, a[0])]
        return w

    def n_constraints(self):
        return sum(1 for op in self.ops if op[0] == "add_constraint")

    def run(self, cols, n_rows, trace_n_rows, coeffs, denom_inv):
        """accumulator columns (from zero) for every row"""
        e, d = n_rows.bit_length() - 1, trace_n_rows.bit_length() - 1
        out = [[0] * n_rows for _ in range(4)]
        for row in range(n_rows):
            reg = {}
            acc = (0, 0, 0, 0)
            k = 0

            def m(r):
                return reg[r][0]

            for op, *a in self.ops:
                if op == "load_const":
                    reg[a[0]] = (a[1], 0, 0, 0)
                elif op == "load_secure_const":
                    reg[a[0]] = tuple(a[1])
                elif op == "load_trace":
                    dst, col, off = a
                    rr = row if off == 0 else neighbor_row(row, d, e, off)
                    reg[dst] = (0 if cols[col] is None else int(cols[col][rr]), 0, 0, 0)
                elif op == "add":
                    reg[a[0]] = ((m(a[1]) + m(a[2])) % P, 0, 0, 0)
                elif op == "sub":
                    reg[a[0]] = ((m(a[1]) - m(a[2])) % P, 0, 0, 0)
                elif op == "mul":
                    reg[a[0]] = (m(a[1]) * m(a[2]) % P, 0, 0, 0)
                elif op == "neg":
                    reg[a[0]] = ((-m(a[1])) % P, 0, 0, 0)
                elif op == "add_const":
                    reg[a[0]] = ((m(a[1]) + a[2]) % P, 0, 0, 0)
                elif op == "mul_const":
                    reg[a[0]] = (m(a[1]) * a[2] % P, 0, 0, 0)
                elif op == "wide_add":
                    reg[a[0]] = pyref.qm_add(reg[a[1]], reg[a[2]])
                elif op == "wide_sub":
                    reg[a[0]] = pyref.qm_sub(reg[a[1]], reg[a[2]])
                elif op == "wide_mul":
                    reg[a[0]] = pyref.qm_mul(reg[a[1]], reg[a[2]])
                elif op == "wide_neg":
                    reg[a[0]] = pyref.qm_sub((0, 0, 0, 0), reg[a[1]])
                elif op == "wide_add_const":
                    reg[a[0]] = pyref.qm_add(reg[a[1]], tuple(a[2]))
                elif op == "wide_mul_const":
                    reg[a[0]] = pyref.qm_mul(reg[a[1]], tuple(a[2]))
                elif op == "wide_add_base":
                    x = reg[a[1]]
                    reg[a[0]] = ((x[0] + m(a[2])) % P, x[1], x[2], x[3])
                elif op == "wide_mul_base":
                    reg[a[0]] = pyref.qm_scale(reg[a[1]], m(a[2]))
                elif op == "base_add_secure_const":
                    reg[a[0]] = pyref.qm_add((m(a[1]), 0, 0, 0), tuple(a[2]))
                elif op == "base_mul_secure_const":
                    reg[a[0]] = pyref.qm_scale(tuple(a[2]), m(a[1]))
                elif op == "widen":
                    reg[a[0]] = (m(a[1]), 0, 0, 0)
                elif op == "combine_ef":
                    reg[a[0]] = (m(a[1]), m(a[2]), m(a[3]), m(a[4]))
                elif op == "add_constraint":
                    acc = pyref.qm_add(acc, pyref.qm_mul(reg[a[0]], tuple(coeffs[k])))
                    k += 1
            res = pyref.qm_scale(acc, int(denom_inv[row >> d]))
            for c in range(4):
                out[c][row] = res[c]
        return out


def domain_points(e):
    """stwo CircleDomain of 2^e points: half_odds(e - 1), then conjugates"""
    hc = pyref.half_odds(e - 1)
    first = [pyref.coset_at(hc, j) for j in range(1 << (e - 1))]
    return first + [(x, (-y) % P) for x, y in first]


_GEOM = {}


def neighbor_row(i, d, e, offset):
    """the bit-reversed row `offset` trace-domain steps from row i"""
    if (d, e) not in _GEOM:
        pts = domain_points(e)
        _GEOM[(d, e)] = (pts, {p: j for j, p in enumerate(pts)}, pyref.gen_pow2(31 - d))
    pts, index, g = _GEOM[(d, e)]
    q = pyref.pt_mul(pts[pyref.bitrev(i, e)], pyref.pt_pow(g, offset % (1 << d)))
    return pyref.bitrev(index[q], e)


def rand_q(rng):
    return tuple(rng.randrange(P) for _ in range(4))


def random_program(rng, n_ops, n_cols, n_regs=40, offsets=(0,), p_constraint=0.1):
    """a random straight-line program; registers are reused and redefined"""
    p = Prog()
    defined = []
    for _ in range(3):
        r = rng.randrange(n_regs)
        p.load_trace(r, rng.randrange(n_cols), rng.choice(offsets))
        defined.append(r)
    binary = ["add", "sub", "mul", "wide_add", "wide_sub", "wide_mul", "wide_add_base", "wide_mul_base"]
    unary = ["neg", "wide_neg", "widen"]
    const1 = ["add_const", "mul_const"]
    constq = ["wide_add_const", "wide_mul_const", "base_add_secure_const", "base_mul_secure_const"]
    for _ in range(n_ops):
        kind = rng.random()
        dst = rng.randrange(n_regs)
        src = lambda: rng.choice(defined)
        if kind < p_constraint:
            p.add_constraint(src())
            continue
        if kind < 0.22:
            p.load_trace(dst, rng.randrange(n_cols), rng.choice(offsets))
        elif kind < 0.26:
            p.load_const(dst, rng.randrange(P))
        elif kind < 0.29:
            p.load_secure_const(dst, rand_q(rng))
        elif kind < 0.62:
            getattr(p, rng.choice(binary))(dst, src(), src())
        elif kind < 0.72:
            getattr(p, rng.choice(unary))(dst, src())
        elif kind < 0.80:
            getattr(p, rng.choice(const1))(dst, src(), rng.randrange(P))
        elif kind < 0.92:
            getattr(p, rng.choice(constq))(dst, src(), rand_q(rng))
        else:
            p.combine_ef(dst, src(), src(), src(), src())
        if dst not in defined:
            defined.append(dst)
    p.add_constraint(rng.choice(defined))
    return p
