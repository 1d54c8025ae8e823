// Circle group over M31 and the coset domains of the prove path
// (src/circle.rs:12-185), shared by kernels and host code.
#pragma once
#include "vx_field.h"

namespace vx {

struct cpt {
  uint32_t x, y;
};

// (x1,y1)*(x2,y2) = (x1x2 - y1y2, x1y2 + y1x2)   (circle.rs:33-38)
VX_HD cpt c_mul(cpt a, cpt b) {
  return {m31_sub(m31_mul(a.x, b.x), m31_mul(a.y, b.y)), m31_add(m31_mul(a.x, b.y), m31_mul(a.y, b.x))};
}
VX_HD cpt c_identity() { return {1u, 0u}; }
// Generator of order 2^31 (circle.rs:26-29).
VX_HD cpt c_generator() { return {2u, 1268011823u}; }
VX_HD cpt c_rdouble(cpt p, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) p = c_mul(p, p);
  return p;
}
VX_HD cpt c_pow(cpt b, uint64_t k) {
  cpt r = c_identity();
  while (k) {
    if (k & 1) r = c_mul(r, b);
    b = c_mul(b, b);
    k >>= 1;
  }
  return r;
}

struct coset {
  cpt init, step;
  uint32_t log_size;
};

enum CosetKind : uint32_t { SUBGROUP = 0, HALF_COSET = 1, HALF_ODDS = 2 };

// circle.rs:101-143. half_odds(30) uses G^1 (saturating 29 - 30).
VX_HD coset make_coset(uint32_t kind, uint32_t L) {
  coset c;
  c.log_size = L;
  c.step = c_rdouble(c_generator(), 31 - L);
  if (kind == SUBGROUP)
    c.init = c_identity();
  else if (kind == HALF_COSET)
    c.init = c_rdouble(c_generator(), 30 - L);
  else
    c.init = c_rdouble(c_generator(), L >= 29 ? 0 : 29 - L);
  return c;
}
// Evaluation domain of the LDE: half_coset(L), or subgroup(31) when L = 31
// (src/prover.rs:655-659, src/verifier.rs:161-165).
VX_HD coset eval_domain(uint32_t L) { return L <= 30 ? make_coset(HALF_COSET, L) : make_coset(SUBGROUP, L); }
VX_HD cpt coset_at(coset c, uint64_t i) { return c_mul(c.init, c_pow(c.step, i)); }

VX_HD uint64_t bitrev64(uint64_t x, uint32_t bits) {
  uint64_t r = 0;
  for (uint32_t i = 0; i < bits; i++) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

}  // namespace vx
