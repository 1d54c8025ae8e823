// M31 / CM31 / QM31 arithmetic shared by the gfx950 kernels and the host.
//
// Semantics follow src/field/m31.rs:24-35,86-113 (canonical values in
// [0, P)), src/field/cm31.rs:76-86 and src/field/qm31.rs:113-132
// (u^2 = 2 + i, word order [a.re, a.im, b.re, b.im]). Arithmetic is exact,
// so the branch-free forms below return the same canonical values as the
// reference's compare-and-subtract forms.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VX_HD __host__ __device__ __forceinline__

namespace vx {

constexpr uint32_t P = 0x7FFFFFFFu;

// r in [0, 2P): unsigned min(r, r - P) selects the canonical representative
// (r - P wraps above 2^31 when r < P).
VX_HD uint32_t m31_canon(uint32_t r) { return r < r - P ? r : r - P; }
VX_HD uint32_t m31_add(uint32_t a, uint32_t b) { return m31_canon(a + b); }
VX_HD uint32_t m31_sub(uint32_t a, uint32_t b) {
  uint32_t r = a - b;
  uint32_t s = r + P;
  return r < s ? r : s;
}
VX_HD uint32_t m31_neg(uint32_t a) { return a == 0 ? 0 : P - a; }
// 62-bit product folded at bit 31 (m31.rs:24-29): x mod P = (x & P) + (x >> 31).
VX_HD uint32_t m31_mul(uint32_t a, uint32_t b) {
  uint64_t x = (uint64_t)a * (uint64_t)b;
  return m31_canon(((uint32_t)x & P) + (uint32_t)(x >> 31));
}
VX_HD uint32_t m31_pow(uint32_t b, uint32_t e) {
  uint32_t r = 1;
  while (e) {
    if (e & 1) r = m31_mul(r, b);
    b = m31_mul(b, b);
    e >>= 1;
  }
  return r;
}
// Fermat inverse (m31.rs:32-35), so inv(0) = 0.
VX_HD uint32_t m31_inv(uint32_t a) { return m31_pow(a, P - 2); }

struct cm31 {
  uint32_t a, b;
};
struct qm31 {
  uint32_t v[4];
};

VX_HD cm31 cm_add(cm31 x, cm31 y) { return {m31_add(x.a, y.a), m31_add(x.b, y.b)}; }
VX_HD cm31 cm_sub(cm31 x, cm31 y) { return {m31_sub(x.a, y.a), m31_sub(x.b, y.b)}; }
VX_HD cm31 cm_mul(cm31 x, cm31 y) {
  return {m31_sub(m31_mul(x.a, y.a), m31_mul(x.b, y.b)),
          m31_add(m31_mul(x.a, y.b), m31_mul(x.b, y.a))};
}
// (p + qi)(2 + i) = (2p - q) + (p + 2q)i
VX_HD cm31 cm_mul_u2(cm31 x) {
  return {m31_sub(m31_add(x.a, x.a), x.b), m31_add(x.a, m31_add(x.b, x.b))};
}
VX_HD cm31 cm_inv(cm31 x) {
  uint32_t in = m31_inv(m31_add(m31_mul(x.a, x.a), m31_mul(x.b, x.b)));
  return {m31_mul(x.a, in), m31_mul(m31_neg(x.b), in)};
}

VX_HD qm31 qm_make(cm31 a, cm31 b) { return {{a.a, a.b, b.a, b.b}}; }
VX_HD qm31 qm_add(qm31 x, qm31 y) {
  return {{m31_add(x.v[0], y.v[0]), m31_add(x.v[1], y.v[1]), m31_add(x.v[2], y.v[2]),
           m31_add(x.v[3], y.v[3])}};
}
VX_HD qm31 qm_sub(qm31 x, qm31 y) {
  return {{m31_sub(x.v[0], y.v[0]), m31_sub(x.v[1], y.v[1]), m31_sub(x.v[2], y.v[2]),
           m31_sub(x.v[3], y.v[3])}};
}
VX_HD qm31 qm_mul(qm31 x, qm31 y) {
  cm31 xa{x.v[0], x.v[1]}, xb{x.v[2], x.v[3]}, ya{y.v[0], y.v[1]}, yb{y.v[2], y.v[3]};
  cm31 ac = cm_mul(xa, ya), bd = cm_mul(xb, yb), ad = cm_mul(xa, yb), bc = cm_mul(xb, ya);
  return qm_make(cm_add(ac, cm_mul_u2(bd)), cm_add(ad, bc));
}
VX_HD qm31 qm_scale(qm31 x, uint32_t s) {
  return {{m31_mul(x.v[0], s), m31_mul(x.v[1], s), m31_mul(x.v[2], s), m31_mul(x.v[3], s)}};
}
VX_HD qm31 qm_inv(qm31 x) {
  cm31 a{x.v[0], x.v[1]}, b{x.v[2], x.v[3]};
  cm31 in = cm_inv(cm_sub(cm_mul(a, a), cm_mul_u2(cm_mul(b, b))));
  cm31 nb{m31_neg(b.a), m31_neg(b.b)};
  return qm_make(cm_mul(a, in), cm_mul(nb, in));
}
VX_HD bool qm_eq(qm31 x, qm31 y) {
  return x.v[0] == y.v[0] && x.v[1] == y.v[1] && x.v[2] == y.v[2] && x.v[3] == y.v[3];
}
// FRI fold of one pair (cuda/fri.cu:58-87; verifier.rs:249-254):
//   (f0 + f1) + alpha * (tw * (f0 - f1))
VX_HD qm31 fri_fold_pair(qm31 f0, qm31 f1, qm31 alpha, uint32_t tw) {
  return qm_add(qm_add(f0, f1), qm_mul(alpha, qm_scale(qm_sub(f0, f1), tw)));
}

// x mod P of any 64-bit x: fold at bit 31 twice, then canonicalise.
VX_HD uint32_t m31_reduce64(uint64_t x) {
  const uint64_t s = (x & P) + (x >> 31);               // < 2^34
  const uint32_t r = (uint32_t)((s & P) + (s >> 31));  // < 2^31 + 8
  return m31_canon(r);
}

// Left multiplication by a fixed QM31 x as a 4x4 matrix over M31 (u^2 =
// 2 + i): (x y)_k = sum_j m[k][j] y_j, from expanding qm_mul:
//   r0 = x0 y0 - x1 y1 + (2 x2 - x3) y2 - (x2 + 2 x3) y3
//   r1 = x1 y0 + x0 y1 + (x2 + 2 x3) y2 + (2 x2 - x3) y3
//   r2 = x2 y0 - x3 y1 + x0 y2 - x1 y3
//   r3 = x3 y0 + x2 y1 + x1 y2 + x0 y3
struct qm31_mat {
  uint32_t m[4][4];
};
VX_HD qm31_mat qm_left_matrix(qm31 x) {
  const uint32_t x0 = x.v[0], x1 = x.v[1], x2 = x.v[2], x3 = x.v[3];
  const uint32_t a = m31_sub(m31_add(x2, x2), x3), b = m31_add(x2, m31_add(x3, x3));
  return {{{x0, m31_neg(x1), a, m31_neg(b)}, {x1, x0, b, a}, {x2, m31_neg(x3), x0, m31_neg(x1)}, {x3, x2, x1, x0}}};
}

// x * y for two row-varying QM31 values through the same matrix: the
// entries need two linear combinations of x and four negations (P - v,
// congruent to -v), each component is four products summed in 64 bits
// (below 4 P^2 < 2^64) and reduced once: about 50 instructions instead of
// the 16 canonical M31 products of qm_mul.
VX_HD qm31 qm_mul_lazy(qm31 x, qm31 y) {
  const uint32_t x0 = x.v[0], x1 = x.v[1], x2 = x.v[2], x3 = x.v[3];
  const uint32_t a = m31_sub(m31_add(x2, x2), x3), b = m31_add(x2, m31_add(x3, x3));
  const uint32_t nx1 = P - x1, nx3 = P - x3, nb = P - b;
  const uint64_t y0 = y.v[0], y1 = y.v[1], y2 = y.v[2], y3 = y.v[3];
  return {{m31_reduce64(x0 * y0 + nx1 * y1 + a * y2 + nb * y3), m31_reduce64(x1 * y0 + x0 * y1 + b * y2 + a * y3),
           m31_reduce64(x2 * y0 + nx3 * y1 + x0 * y2 + nx1 * y3), m31_reduce64(x3 * y0 + x2 * y1 + x1 * y2 + x0 * y3)}};
}

// The same fold with alpha given as its matrix: each component is a dot
// product of four M31 products accumulated in 64 bits (4 P^2 < 2^64) and
// reduced once, then tw * s + (f0 + f1), again reduced once.
VX_HD qm31 fri_fold_pair_m(qm31 f0, qm31 f1, const qm31_mat& am, uint32_t tw) {
  uint32_t d[4];
  for (int j = 0; j < 4; j++) d[j] = m31_sub(f0.v[j], f1.v[j]);
  qm31 r;
  for (int k = 0; k < 4; k++) {
    uint64_t acc = 0;
    for (int j = 0; j < 4; j++) acc += (uint64_t)am.m[k][j] * d[j];
    const uint32_t s = m31_reduce64(acc);
    r.v[k] = m31_reduce64((uint64_t)tw * s + f0.v[k] + f1.v[k]);
  }
  return r;
}

}  // namespace vx
