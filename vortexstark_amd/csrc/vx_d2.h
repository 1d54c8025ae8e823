// Synthetic degree-2 multi-column AIR: the serving workload's 34-column,
// 35-constraint request type. It has no counterpart in the reference; its
// shape mirrors the reference's 34-column Cairo trace (trace columns in
// Merkle trees of <= 16 columns, src/cairo_air/prover.rs:341-350, since one
// leaf hashes at most 16 words, cuda/blake2s.cu:48-53; constraints checked at
// the query rows from values at q and q + 1, prover.rs:415-419).
//
//   trace     t_c[i] = mix64(seed + G ((c << 32 | i) + 1)) mod P,
//             G = 0x9E3779B97F4A7C15, mix64 = the SplitMix64 finaliser
//   AIR       C_k(i) = c_k t_a[i] t_b[i+1] + d_k t_e[i] for k < K, with
//             a = k mod W, b = (5k + 1) mod W, e = (11k + 7) mod W and
//             (c_0, d_0, c_1, d_1, ...) = SplitMix64(air_seed) outputs mod P
//   quotient  Q(i) = sum_k alpha^(k+1) C_k(i) over the LDE (row i + 1 taken
//             mod N in evaluation order, as the Fibonacci quotient does,
//             cuda/constraint.cu:22-30)
// Shared by the kernels, the host prover and the host verifier.
#pragma once
#include "vx_field.h"

namespace vx {

constexpr uint64_t D2_GOLDEN = 0x9E3779B97F4A7C15ull;
constexpr uint32_t D2_GROUP = 16;   // columns per trace tree
constexpr uint32_t D2_MAX_COLS = 256;
constexpr uint32_t D2_MAX_CONS = 1024;

VX_HD uint64_t d2_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// x mod (2^31 - 1) of a 64-bit x: fold at bit 31 twice, then canonicalise.
VX_HD uint32_t d2_mod_p(uint64_t x) { return m31_reduce64(x); }
VX_HD uint32_t d2_value(uint64_t seed, uint32_t col, uint64_t row) {
  return d2_mod_p(d2_mix64(seed + D2_GOLDEN * ((((uint64_t)col << 32) | row) + 1)));
}
VX_HD uint32_t d2_col_a(uint32_t k, uint32_t w) { return k % w; }
VX_HD uint32_t d2_col_b(uint32_t k, uint32_t w) { return (5 * k + 1) % w; }
VX_HD uint32_t d2_col_e(uint32_t k, uint32_t w) { return (11 * k + 7) % w; }
VX_HD uint32_t d2_groups(uint32_t w) { return (w + D2_GROUP - 1) / D2_GROUP; }
VX_HD uint32_t d2_group_cols(uint32_t w, uint32_t g) {
  return w - D2_GROUP * g < D2_GROUP ? w - D2_GROUP * g : D2_GROUP;
}
// (c_0, d_0, c_1, d_1, ...) of the K constraints.
inline void d2_coeffs(uint64_t air_seed, uint32_t k_cons, uint32_t* cd) {
  uint64_t s = air_seed;
  for (uint32_t k = 0; k < 2 * k_cons; k++) {
    s += D2_GOLDEN;
    cd[k] = d2_mod_p(d2_mix64(s));
  }
}

}  // namespace vx
