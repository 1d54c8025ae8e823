// Cairo VM AIR constraint quotient of one row (cuda/cairo_constraint.cu:
// 157-580 and the chunked form :582-850), shared by the stand-alone kernels
// (cairo.hip) and the fused quotient + Merkle leaf producer (merkle.hip).
//
// Columns (cuda/cairo_constraint.cu:8-53): pc, ap, fp, inst_lo, inst_hi,
// 15 flags, dst_addr, dst, op0_addr, op0, op1_addr, op1, res, off0..2,
// dst_inv, dict key / new / active. Interaction columns are QM31 SoA groups
// of 4: S_logup, T1, T2, T3, S_rc, U1, U2, S_dict (the chunked form reads
// S_logup, S_rc, S_dict). Challenges: z_mem, alpha_mem, alpha_mem^2, z_rc,
// z_dict, alpha_dict.
//
// Q(row) = (sum_k alpha_k c_k(row)) / Z_H(row). The alpha-weighted sum is
// accumulated in four 64-bit lanes without reduction: an M31 constraint
// adds alpha_k * c (four products below P^2), a QM31 constraint adds
// M(alpha_k) c (the 4 x 4 matrix of left multiplication, four products per
// lane); lanes are folded to lo + 2 hi (congruent mod P, below 3 * 2^32)
// often enough that no sum reaches 2^64, and reduced once per row.
#pragma once
#include "vx_internal.h"

namespace vx {

enum : uint32_t {
  CA_PC = 0, CA_AP = 1, CA_FP = 2, CA_INST_LO = 3, CA_INST_HI = 4, CA_FLAGS = 5, CA_DST_ADDR = 20, CA_DST = 21,
  CA_OP0_ADDR = 22, CA_OP0 = 23, CA_OP1_ADDR = 24, CA_OP1 = 25, CA_RES = 26, CA_OFF0 = 27, CA_OFF1 = 28,
  CA_OFF2 = 29, CA_DST_INV = 30, CA_DICT_KEY = 31, CA_DICT_NEW = 32, CA_DICT_ACTIVE = 33, CA_COLS = 34
};
constexpr uint32_t CA_BIAS = 0x8000u;

// Row r (0 <= r < rows) of a launch reads columns at col_base + r, the next
// row at (that + blowup_step) mod wrap (no wrap when wrap = 0), Z_H^-1 and
// the transition factor at g_base + r, and writes Q at o_base + r:
//   full   col_base 0,   wrap n, g_base 0,   o_base 0
//   slab   col_base 0,   wrap 0, g_base off, o_base off  (slabs of rows + step)
//   chunk  col_base off, wrap N, g_base off, o_base 0    (chunked form)
struct CairoArgs {
  const uint32_t* col[CA_COLS];
  const uint32_t* inter[32];
  uint32_t* out[4];
  const uint32_t* alpha;   // 4 words per weighted constraint (40, or 35 in the chunked form)
  const uint32_t* vh_inv;
  const uint32_t* trans;
  const uint32_t* chal;    // 24 words
  uint64_t col_base, wrap, g_base, o_base;
  uint32_t blowup_step;
  uint32_t chunked;        // 1: the chunked kernel's 35-constraint form
};

__device__ __forceinline__ uint64_t ca_fold(uint64_t x) { return (x & 0xFFFFFFFFull) + ((x >> 32) << 1); }

struct CaAcc {
  uint64_t a[4];
};

// acc += alpha * c for an M31 c (acc below 2^63 before, 4 P^2 headroom kept
// by folding every third call via ca_m31_fold)
template <typename A>
__device__ __forceinline__ void ca_add_m31(CaAcc& acc, const A& args, uint32_t k, uint32_t c) {
  const uint4 al = ld_uniform(reinterpret_cast<const uint4*>(args.alpha + 4 * k));
  acc.a[0] += (uint64_t)al.x * c;
  acc.a[1] += (uint64_t)al.y * c;
  acc.a[2] += (uint64_t)al.z * c;
  acc.a[3] += (uint64_t)al.w * c;
}
__device__ __forceinline__ void ca_fold_all(CaAcc& acc) {
#pragma unroll
  for (int j = 0; j < 4; j++) acc.a[j] = ca_fold(acc.a[j]);
}
// acc += alpha * c for a QM31 c: the left-multiplication matrix of alpha
// (uniform, so built once per wave) times c, four products per lane.
template <typename A>
__device__ __forceinline__ void ca_add_qm31(CaAcc& acc, const A& args, uint32_t k, const qm31& c) {
  const uint4 al = ld_uniform(reinterpret_cast<const uint4*>(args.alpha + 4 * k));
  const qm31_mat m = qm_left_matrix(qm31{{al.x, al.y, al.z, al.w}});
  ca_fold_all(acc);
#pragma unroll
  for (int j = 0; j < 4; j++)
    acc.a[j] += (uint64_t)m.m[j][0] * c.v[0] + (uint64_t)m.m[j][1] * c.v[1] + (uint64_t)m.m[j][2] * c.v[2] +
                (uint64_t)m.m[j][3] * c.v[3];
  ca_fold_all(acc);
}

__device__ __forceinline__ qm31 ca_lift(uint32_t a) { return qm31{{a, 0, 0, 0}}; }
template <typename A>
__device__ __forceinline__ qm31 ca_chal(const A& args, int k) {
  const uint4 v = ld_uniform(reinterpret_cast<const uint4*>(args.chal + 4 * k));
  return qm31{{v.x, v.y, v.z, v.w}};
}
template <typename A>
__device__ __forceinline__ qm31 ca_inter(const A& args, int g, uint64_t i) {
  return qm31{{as_global(args.inter[4 * g + 0])[i], as_global(args.inter[4 * g + 1])[i],
               as_global(args.inter[4 * g + 2])[i], as_global(args.inter[4 * g + 3])[i]}};
}
template <typename A>
__device__ __forceinline__ uint32_t ca_col(const A& args, uint32_t c, uint64_t i) {
  return as_global(args.col[c])[i];
}

// Q of launch row r; `args` is the kernel-argument struct or a uniform
// (constant-address) view of one in device memory.
template <typename A>
__device__ __forceinline__ qm31 cairo_row(const A& args, uint64_t r) {
  const uint64_t i = args.col_base + r;
  // i < wrap and blowup_step <= wrap (checked on the host), so one
  // conditional subtraction wraps the next row
  uint64_t ni = i + args.blowup_step;
  if (args.wrap && ni >= args.wrap) ni -= args.wrap;
  const uint64_t g = args.g_base + r;
  const uint32_t tf = as_global(args.trans)[g];
  uint32_t f[15];
#pragma unroll
  for (int j = 0; j < 15; j++) f[j] = ca_col(args, CA_FLAGS + j, i);
  const uint32_t pc = ca_col(args, CA_PC, i), ap = ca_col(args, CA_AP, i), fp = ca_col(args, CA_FP, i);
  const uint32_t dst = ca_col(args, CA_DST, i), op0 = ca_col(args, CA_OP0, i), op1 = ca_col(args, CA_OP1, i);
  const uint32_t res = ca_col(args, CA_RES, i);
  const uint32_t npc = ca_col(args, CA_PC, ni), nap = ca_col(args, CA_AP, ni), nfp = ca_col(args, CA_FP, ni);
  const uint32_t off0 = ca_col(args, CA_OFF0, i), off1 = ca_col(args, CA_OFF1, i), off2 = ca_col(args, CA_OFF2, i);
  const uint32_t op1_imm = f[2], res_add = f[5], res_mul = f[6], jabs = f[7], jrel = f[8], jnz = f[9];
  const uint32_t ap_add = f[10], ap_add1 = f[11], call = f[12], ret = f[13], asrt = f[14];
  CaAcc acc{{0, 0, 0, 0}};
  uint32_t k = 0;
  // 0..14 flags are binary
#pragma unroll
  for (int j = 0; j < 15; j++) {
    ca_add_m31(acc, args, k++, m31_mul(f[j], m31_sub(1, f[j])));
    if (j % 3 == 2) ca_fold_all(acc);
  }
  {  // 15 result
    const uint32_t e = m31_add(m31_add(m31_mul(m31_sub(m31_sub(1, res_add), res_mul), op1),
                                       m31_mul(res_add, m31_add(op0, op1))),
                               m31_mul(res_mul, m31_mul(op0, op1)));
    ca_add_m31(acc, args, k++, m31_mul(m31_sub(1, jnz), m31_sub(res, e)));
  }
  const uint32_t inst_size = m31_add(1, op1_imm);
  {  // 16 pc update
    const uint32_t notj = m31_sub(m31_sub(m31_sub(1, jabs), jrel), jnz);
    const uint32_t tgt = m31_add(m31_add(m31_mul(notj, m31_add(pc, inst_size)), m31_mul(jabs, res)),
                                 m31_mul(jrel, m31_add(pc, res)));
    const uint32_t nonjnz = m31_mul(m31_sub(1, jnz), m31_sub(npc, tgt));
    const uint32_t jz = m31_mul(jnz, m31_mul(dst, m31_sub(npc, m31_add(pc, op1))));
    ca_add_m31(acc, args, k++, m31_mul(m31_add(nonjnz, jz), tf));
  }
  // 17 ap update
  ca_add_m31(acc, args, k++,
             m31_mul(m31_sub(nap, m31_add(m31_add(m31_add(ap, m31_mul(ap_add, res)), ap_add1), m31_mul(call, 2))),
                     tf));
  ca_fold_all(acc);
  // 18 fp update, 19 assert_eq
  ca_add_m31(acc, args, k++,
             m31_mul(m31_sub(nfp, m31_add(m31_add(m31_mul(m31_sub(m31_sub(1, call), ret), fp),
                                                  m31_mul(call, m31_add(ap, 2))),
                                          m31_mul(ret, dst))),
                     tf));
  ca_add_m31(acc, args, k++, m31_mul(asrt, m31_sub(dst, res)));
  // 20..22 operand addresses (offsets biased by 2^15)
  ca_add_m31(acc, args, k++,
             m31_sub(ca_col(args, CA_DST_ADDR, i),
                     m31_sub(m31_add(m31_add(m31_mul(m31_sub(1, f[0]), ap), m31_mul(f[0], fp)), off0), CA_BIAS)));
  ca_fold_all(acc);
  ca_add_m31(acc, args, k++,
             m31_sub(ca_col(args, CA_OP0_ADDR, i),
                     m31_sub(m31_add(m31_add(m31_mul(m31_sub(1, f[1]), ap), m31_mul(f[1], fp)), off1), CA_BIAS)));
  {
    const uint32_t dflt = m31_sub(m31_sub(m31_sub(1, op1_imm), f[3]), f[4]);
    const uint32_t base = m31_add(m31_add(m31_add(m31_mul(op1_imm, pc), m31_mul(f[3], fp)), m31_mul(f[4], ap)),
                                  m31_mul(dflt, op0));
    ca_add_m31(acc, args, k++, m31_sub(ca_col(args, CA_OP1_ADDR, i), m31_sub(m31_add(base, off2), CA_BIAS)));
  }
  const uint32_t dinv = ca_col(args, CA_DST_INV, i);
  const uint32_t one_m_ddi = m31_sub(1, m31_mul(dst, dinv));
  // 23 jnz fall-through, 24 jnz inverse
  ca_add_m31(acc, args, k++,
             m31_mul(m31_mul(jnz, m31_mul(one_m_ddi, m31_sub(npc, m31_add(pc, inst_size)))), tf));
  ca_fold_all(acc);
  ca_add_m31(acc, args, k++, m31_mul(jnz, m31_mul(dst, one_m_ddi)));
  // 25..29 exclusivity
  ca_add_m31(acc, args, k++, m31_mul(op1_imm, f[3]));
  ca_add_m31(acc, args, k++, m31_mul(op1_imm, f[4]));
  ca_fold_all(acc);
  ca_add_m31(acc, args, k++, m31_mul(f[3], f[4]));
  ca_add_m31(acc, args, k++, m31_add(m31_add(m31_mul(jabs, jrel), m31_mul(jabs, jnz)), m31_mul(jrel, jnz)));
  ca_add_m31(acc, args, k++, m31_add(m31_add(m31_mul(call, ret), m31_mul(call, asrt)), m31_mul(ret, asrt)));
  ca_fold_all(acc);
  {  // 30 instruction word: 2^31 = 1, 2^32 = 2, 2^(48+j) = 2^(17+j), 2^62 = 1 mod P
    uint32_t rhs = m31_add(m31_add(off0, m31_mul(off1, 1u << 16)), m31_mul(off2, 2));
#pragma unroll
    for (int j = 0; j < 14; j++) rhs = m31_add(rhs, m31_mul(f[j], 1u << (17 + j)));
    rhs = m31_add(rhs, f[14]);
    ca_add_m31(acc, args, k++, m31_sub(m31_add(ca_col(args, CA_INST_LO, i), ca_col(args, CA_INST_HI, i)), rhs));
  }
  // memory (LogUp) denominators d = z - (addr + alpha value) (the
  // instruction fetch adds alpha^2 inst_hi): each component is z_k plus
  // products with the negated uniform alphas, summed in 64 bits and
  // reduced once
  const qm31 z_mem = ca_chal(args, 0), a_mem = ca_chal(args, 1), z_rc = ca_chal(args, 3);
  const qm31 a_mem2 = ca_chal(args, 2);
  const uint32_t inst_lo = ca_col(args, CA_INST_LO, i), inst_hi = ca_col(args, CA_INST_HI, i);
  auto denom = [&](uint32_t addr, uint32_t val) {
    qm31 d;
#pragma unroll
    for (int c = 0; c < 4; c++)
      d.v[c] = m31_reduce64((uint64_t)z_mem.v[c] + (c == 0 ? (uint64_t)(P - addr) : 0) +
                            (uint64_t)(P - a_mem.v[c]) * val);
    return d;
  };
  qm31 d0;
#pragma unroll
  for (int c = 0; c < 4; c++)
    d0.v[c] = m31_reduce64((uint64_t)z_mem.v[c] + (c == 0 ? (uint64_t)(P - pc) : 0) +
                           (uint64_t)(P - a_mem.v[c]) * inst_lo + (uint64_t)(P - a_mem2.v[c]) * inst_hi);
  const qm31 d1 = denom(ca_col(args, CA_DST_ADDR, i), dst);
  const qm31 d2 = denom(ca_col(args, CA_OP0_ADDR, i), op0);
  const qm31 d3 = denom(ca_col(args, CA_OP1_ADDR, i), op1);
  const qm31 r0 = qm_sub(z_rc, ca_lift(off0)), r1 = qm_sub(z_rc, ca_lift(off1)), r2 = qm_sub(z_rc, ca_lift(off2));
  const qm31 one{{1, 0, 0, 0}};
  if (!args.chunked) {
    // 31a..31d, 32a..32c in polynomial form over T1..T3, U1, U2
    const qm31 s = ca_inter(args, 0, i), t1 = ca_inter(args, 1, i), t2 = ca_inter(args, 2, i),
               t3 = ca_inter(args, 3, i);
    ca_add_qm31(acc, args, k++, qm_sub(qm_mul_lazy(qm_sub(t1, s), d0), one));
    ca_add_qm31(acc, args, k++, qm_sub(qm_mul_lazy(qm_sub(t2, t1), d1), one));
    ca_add_qm31(acc, args, k++, qm_sub(qm_mul_lazy(qm_sub(t3, t2), d2), one));
    ca_add_qm31(acc, args, k++, qm_scale(qm_sub(qm_mul_lazy(qm_sub(ca_inter(args, 0, ni), t3), d3), one), tf));
    const qm31 rs = ca_inter(args, 4, i), u1 = ca_inter(args, 5, i), u2 = ca_inter(args, 6, i);
    ca_add_qm31(acc, args, k++, qm_sub(qm_mul_lazy(qm_sub(u1, rs), r0), one));
    ca_add_qm31(acc, args, k++, qm_sub(qm_mul_lazy(qm_sub(u2, u1), r1), one));
    ca_add_qm31(acc, args, k++, qm_scale(qm_sub(qm_mul_lazy(qm_sub(ca_inter(args, 4, ni), u2), r2), one), tf));
  } else {
    // 31, 32 as one step each: S_next - S - sum of inverse denominators
    const qm31 dl = qm_add(qm_add(qm_add(qm_inv(d0), qm_inv(d1)), qm_inv(d2)), qm_inv(d3));
    ca_add_qm31(acc, args, k++, qm_scale(qm_sub(qm_sub(ca_inter(args, 0, ni), ca_inter(args, 0, i)), dl), tf));
    const qm31 dr = qm_add(qm_add(qm_inv(r0), qm_inv(r1)), qm_inv(r2));
    ca_add_qm31(acc, args, k++, qm_scale(qm_sub(qm_sub(ca_inter(args, 1, ni), ca_inter(args, 1, i)), dr), tf));
  }
  // 33 dict_active binary, 34 dict step: (S_next - S)(z_dict - entry) - active
  const uint32_t act = ca_col(args, CA_DICT_ACTIVE, i);
  ca_add_m31(acc, args, k++, m31_mul(act, m31_sub(1, act)));
  const int sd = args.chunked ? 2 : 7;
  // z_dict - (key + alpha_dict new), as the memory denominators
  const qm31 z_d = ca_chal(args, 4), a_d = ca_chal(args, 5);
  const uint32_t key = ca_col(args, CA_DICT_KEY, i), nv = ca_col(args, CA_DICT_NEW, i);
  qm31 dd;
#pragma unroll
  for (int c = 0; c < 4; c++)
    dd.v[c] = m31_reduce64((uint64_t)z_d.v[c] + (c == 0 ? (uint64_t)(P - key) : 0) + (uint64_t)(P - a_d.v[c]) * nv);
  const qm31 c34 =
      qm_scale(qm_sub(qm_mul_lazy(qm_sub(ca_inter(args, sd, ni), ca_inter(args, sd, i)), dd), ca_lift(act)), tf);
  ca_add_qm31(acc, args, k++, c34);
  const uint32_t vh = as_global(args.vh_inv)[g];
  return qm31{{m31_mul(m31_reduce64(acc.a[0]), vh), m31_mul(m31_reduce64(acc.a[1]), vh),
               m31_mul(m31_reduce64(acc.a[2]), vh), m31_mul(m31_reduce64(acc.a[3]), vh)}};
}

}  // namespace vx
