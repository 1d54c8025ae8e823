"""Independent Python restatement of the Cairo VM AIR quotient, written from
the reference's host-side evaluation of the same constraints at query points
(src/cairo_air/prover.rs:4294-4580, the verifier's check of
Q(x) * Z_H = C(x)) and the LogUp / range-check step form of the chunked
kernel (cuda/cairo_constraint.cu:120-153). Not derived from oracle/vx_oracle.c.

Also a generator of VALID Cairo rows: every constraint holds, so the
quotient must be zero wherever the transition is enforced.

Columns (cuda/cairo_constraint.cu:8-53): 0 pc, 1 ap, 2 fp, 3 inst_lo,
4 inst_hi, 5..19 flags (dst_reg, op0_reg, op1_imm, op1_fp, op1_ap, res_add,
res_mul, pc_jump_abs, pc_jump_rel, pc_jnz, ap_add, ap_add1, call, ret,
assert), 20 dst_addr, 21 dst, 22 op0_addr, 23 op0, 24 op1_addr, 25 op1,
26 res, 27..29 off0..2, 30 dst_inv, 31 dict_key, 32 dict_new, 33 dict_active.
Interaction columns (QM31, one tuple per row): S_logup, T1, T2, T3, S_rc,
U1, U2, S_dict. Challenges: z_mem, alpha_mem, alpha_mem^2, z_rc, z_dict,
alpha_dict.
"""
from pyref import P, qm_add, qm_mul, qm_scale, qm_sub

ONE = (1, 0, 0, 0)
BIAS = 1 << 15


def lift(a):
    return (a % P, 0, 0, 0)


def qm_inv(x):
    # x^(|QM31| - 2) by square-and-multiply; 0 -> 0
    e, r, b = P ** 4 - 2, ONE, x
    while e:
        if e & 1:
            r = qm_mul(r, b)
        b = qm_mul(b, b)
        e >>= 1
    return r


def base_constraints(row, nxt, tf):
    """Constraints 0..30 (prover.rs:4297-4483) as M31 values, in order."""
    m = lambda v: v % P  # noqa: E731
    f = [row[5 + j] for j in range(15)]
    pc, ap, fp = row[0], row[1], row[2]
    dst, op0, op1, res = row[21], row[23], row[25], row[26]
    npc, nap, nfp = nxt[0], nxt[1], nxt[2]
    op1_imm, res_add, res_mul = f[2], f[5], f[6]
    jabs, jrel, jnz = f[7], f[8], f[9]
    ap_add, ap_add1, call, ret, asrt = f[10], f[11], f[12], f[13], f[14]
    out = [m(x * (1 - x)) for x in f]
    expected = (1 - res_add - res_mul) * op1 + res_add * (op0 + op1) + res_mul * op0 * op1
    out.append(m((1 - jnz) * (res - expected)))
    inst_size = 1 + op1_imm
    regular = (1 - jabs - jrel - jnz) * (pc + inst_size)
    non_jnz = (1 - jnz) * (npc - (regular + jabs * res + jrel * (pc + res)))
    jnz_part = jnz * dst * (npc - (pc + op1))
    out.append(m((non_jnz + jnz_part) * tf))
    out.append(m((nap - (ap + ap_add * res + ap_add1 + 2 * call)) * tf))
    out.append(m((nfp - ((1 - call - ret) * fp + call * (ap + 2) + ret * dst)) * tf))
    out.append(m(asrt * (dst - res)))
    off0, off1, off2 = row[27], row[28], row[29]
    out.append(m(row[20] - ((1 - f[0]) * ap + f[0] * fp + off0 - BIAS)))
    out.append(m(row[22] - ((1 - f[1]) * ap + f[1] * fp + off1 - BIAS)))
    base = op1_imm * pc + f[3] * fp + f[4] * ap + (1 - op1_imm - f[3] - f[4]) * op0
    out.append(m(row[24] - (base + off2 - BIAS)))
    dinv = row[30]
    out.append(m(jnz * (1 - dst * dinv) * (npc - pc - inst_size) * tf))
    out.append(m(jnz * dst * (1 - dst * dinv)))
    out.append(m(op1_imm * f[3]))
    out.append(m(op1_imm * f[4]))
    out.append(m(f[3] * f[4]))
    out.append(m(jabs * jrel + jabs * jnz + jrel * jnz))
    out.append(m(call * ret + call * asrt + ret * asrt))
    rhs = off0 + off1 * (1 << 16) + off2 * 2 + sum(f[j] * (1 << (17 + j)) for j in range(14)) + f[14]
    out.append(m(row[3] + row[4] - rhs))
    return out


def denominators(row, chal):
    z, a, a2 = chal[0], chal[1], chal[2]
    d0 = qm_sub(z, qm_add(lift(row[0]), qm_add(qm_scale(a, row[3]), qm_scale(a2, row[4]))))
    d1 = qm_sub(z, qm_add(lift(row[20]), qm_scale(a, row[21])))
    d2 = qm_sub(z, qm_add(lift(row[22]), qm_scale(a, row[23])))
    d3 = qm_sub(z, qm_add(lift(row[24]), qm_scale(a, row[25])))
    return d0, d1, d2, d3


def dict_denominator(row, chal):
    return qm_sub(chal[4], qm_add(lift(row[31]), qm_scale(chal[5], row[32])))


def quotient_row(row, nxt, it, it_next, alphas, chal, tf, vh_inv):
    """Q at one row, 40 weighted constraints (prover.rs:4294-4580)."""
    cs = [lift(c) for c in base_constraints(row, nxt, tf)]
    s, t1, t2, t3, r, u1, u2, sd = it
    d0, d1, d2, d3 = denominators(row, chal)
    cs.append(qm_sub(qm_mul(qm_sub(t1, s), d0), ONE))
    cs.append(qm_sub(qm_mul(qm_sub(t2, t1), d1), ONE))
    cs.append(qm_sub(qm_mul(qm_sub(t3, t2), d2), ONE))
    cs.append(qm_scale(qm_sub(qm_mul(qm_sub(it_next[0], t3), d3), ONE), tf))
    r0, r1, r2 = (qm_sub(chal[3], lift(row[27 + k])) for k in range(3))
    cs.append(qm_sub(qm_mul(qm_sub(u1, r), r0), ONE))
    cs.append(qm_sub(qm_mul(qm_sub(u2, u1), r1), ONE))
    cs.append(qm_scale(qm_sub(qm_mul(qm_sub(it_next[4], u2), r2), ONE), tf))
    act = row[33]
    cs.append(lift(act * (1 - act)))
    cs.append(qm_scale(qm_sub(qm_mul(qm_sub(it_next[7], sd), dict_denominator(row, chal)), lift(act)), tf))
    assert len(cs) == 40
    acc = (0, 0, 0, 0)
    for a, c in zip(alphas, cs):
        acc = qm_add(acc, qm_mul(a, c))
    return qm_scale(acc, vh_inv)


def quotient_row_chunk(row, nxt, s3, s3_next, alphas, chal, tf, vh_inv):
    """Q at one row of the chunked kernel: 35 weighted constraints, the
    LogUp and range-check steps as S_next - S - sum of inverse denominators
    (cuda/cairo_constraint.cu:120-153, 785-810)."""
    cs = [lift(c) for c in base_constraints(row, nxt, tf)]
    delta = (0, 0, 0, 0)
    for d in denominators(row, chal):
        delta = qm_add(delta, qm_inv(d))
    cs.append(qm_scale(qm_sub(qm_sub(s3_next[0], s3[0]), delta), tf))
    delta_rc = (0, 0, 0, 0)
    for k in range(3):
        delta_rc = qm_add(delta_rc, qm_inv(qm_sub(chal[3], lift(row[27 + k]))))
    cs.append(qm_scale(qm_sub(qm_sub(s3_next[1], s3[1]), delta_rc), tf))
    act = row[33]
    cs.append(lift(act * (1 - act)))
    cs.append(qm_scale(qm_sub(qm_mul(qm_sub(s3_next[2], s3[2]), dict_denominator(row, chal)), lift(act)), tf))
    assert len(cs) == 35
    acc = (0, 0, 0, 0)
    for a, c in zip(alphas, cs):
        acc = qm_add(acc, qm_mul(a, c))
    return qm_scale(acc, vh_inv)


# ---------------- valid rows ----------------

def valid_step(rng, pc, ap, fp, s_logup, s_rc, s_dict, chal):
    """One valid Cairo step from registers (pc, ap, fp) and running sums:
    returns (row[34], interaction values[8], next registers, next sums).
    Flags are drawn so that every exclusivity holds; values derive from them."""
    f = [0] * 15
    f[0], f[1] = rng.randrange(2), rng.randrange(2)              # dst_reg, op0_reg
    op1_src = rng.randrange(4)                                    # [op0 + off2], imm, fp, ap
    f[2], f[3], f[4] = int(op1_src == 1), int(op1_src == 2), int(op1_src == 3)
    res_kind = rng.randrange(3)                                   # default, add, mul
    f[5], f[6] = int(res_kind == 1), int(res_kind == 2)
    pc_kind = rng.randrange(4)                                    # regular, abs, rel, jnz
    f[7], f[8], f[9] = int(pc_kind == 1), int(pc_kind == 2), int(pc_kind == 3)
    f[10], f[11] = rng.randrange(2), rng.randrange(2)
    op_kind = rng.randrange(4)                                    # none, call, ret, assert
    f[12], f[13], f[14] = int(op_kind == 1), int(op_kind == 2), int(op_kind == 3)
    off = [rng.randrange(1 << 16) for _ in range(3)]
    op0 = rng.randrange(P)
    op1 = rng.randrange(P)
    if f[9]:   # jnz: the result constraint is off; dst zero (fall through) or not (jump)
        dst = rng.choice([0, rng.randrange(1, P)])
        res = dst if f[14] else rng.randrange(P)
    else:
        res = ((1 - f[5] - f[6]) * op1 + f[5] * (op0 + op1) + f[6] * op0 * op1) % P
        dst = res if f[14] else rng.randrange(P)
    dst_inv = pow(dst, P - 2, P) if dst else rng.randrange(P)
    inst_size = 1 + f[2]
    if f[9]:
        npc = (pc + op1) % P if dst else (pc + inst_size) % P
    else:
        npc = ((1 - f[7] - f[8]) * (pc + inst_size) + f[7] * res + f[8] * (pc + res)) % P
    nap = (ap + f[10] * res + f[11] + 2 * f[12]) % P
    nfp = ((1 - f[12] - f[13]) * fp + f[12] * (ap + 2) + f[13] * dst) % P
    dst_addr = ((1 - f[0]) * ap + f[0] * fp + off[0] - BIAS) % P
    op0_addr = ((1 - f[1]) * ap + f[1] * fp + off[1] - BIAS) % P
    op1_addr = (f[2] * pc + f[3] * fp + f[4] * ap + (1 - f[2] - f[3] - f[4]) * op0 + off[2] - BIAS) % P
    rhs = (off[0] + off[1] * (1 << 16) + off[2] * 2 + sum(f[j] * (1 << (17 + j)) for j in range(14)) + f[14]) % P
    inst_hi = rng.randrange(P)
    inst_lo = (rhs - inst_hi) % P
    key, new, act = rng.randrange(P), rng.randrange(P), rng.randrange(2)
    row = [pc, ap, fp, inst_lo, inst_hi] + f + [dst_addr, dst, op0_addr, op0, op1_addr, op1, res] + off + \
          [dst_inv, key, new, act]
    d0, d1, d2, d3 = denominators(row, chal)
    t1 = qm_add(s_logup, qm_inv(d0))
    t2 = qm_add(t1, qm_inv(d1))
    t3 = qm_add(t2, qm_inv(d2))
    s_next = qm_add(t3, qm_inv(d3))
    r0, r1, r2 = (qm_sub(chal[3], lift(off[k])) for k in range(3))
    u1 = qm_add(s_rc, qm_inv(r0))
    u2 = qm_add(u1, qm_inv(r1))
    rc_next = qm_add(u2, qm_inv(r2))
    sd_next = qm_add(s_dict, qm_scale(qm_inv(dict_denominator(row, chal)), act))
    inter = [s_logup, t1, t2, t3, s_rc, u1, u2, s_dict]
    return row, inter, (npc, nap, nfp), (s_next, rc_next, sd_next)


def valid_table(rng, n, blowup_step, chal):
    """n rows in which row i -> row (i + blowup_step) mod n is a valid step
    along each of the blowup_step cycles; the step from the last row of a
    cycle back to its first is not (it needs trans = 0 there). Returns
    (rows, inter, last) with last[i] = True for those cycle ends."""
    rows, inter = [None] * n, [None] * n
    last = [False] * n
    for start in range(blowup_step):
        regs = (rng.randrange(P), rng.randrange(P), rng.randrange(P))
        sums = tuple(tuple(rng.randrange(P) for _ in range(4)) for _ in range(3))
        i = start
        while True:
            # the row at i starts from the previous step's registers and sums
            rows[i], inter[i], regs, sums = valid_step(rng, *regs, *sums, chal)
            i_next = (i + blowup_step) % n
            if i_next == start:
                last[i] = True
                break
            i = i_next
    return rows, inter, last
