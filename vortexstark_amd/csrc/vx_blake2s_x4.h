// Quad-parallel Blake2s (device only): the four lanes of a quad compute one
// compression together, lane q holding column q of the state (v[q], v[4+q],
// v[8+q], v[12+q]). The column step is one G per lane; for the diagonal step
// lane q needs v[4+(q+1)%4], v[8+(q+2)%4], v[12+(q+3)%4], which DPP quad
// permutes bring in (and take back) register to register. A compression is
// then ~440 VALU instructions per lane instead of 960, so a dependent chain of
// node hashes (the top levels of a tree, where few nodes are left and most
// SIMDs idle) runs about twice as fast. Same function as b2_compress.
#pragma once
#include "vx_blake2s.h"

namespace vx {

// RFC 7693 message schedule.
__device__ constexpr uint8_t B2_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};

// quad_perm controls: lane q reads lane (q + k) % 4
#define VX_DPP_ROT1 0x39  // [1,2,3,0]
#define VX_DPP_ROT2 0x4E  // [2,3,0,1]
#define VX_DPP_ROT3 0x93  // [3,0,1,2]

__device__ __forceinline__ uint32_t quad_rot(uint32_t x, int ctrl) {
  // ctrl must be a compile-time constant at every call site
  switch (ctrl) {
    case VX_DPP_ROT1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, VX_DPP_ROT1, 0xF, 0xF, true);
    case VX_DPP_ROT2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, VX_DPP_ROT2, 0xF, 0xF, true);
    default: return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, VX_DPP_ROT3, 0xF, 0xF, true);
  }
}

// x = word k of {A, B, C, D} by lane q (two bits -> three selects)
__device__ __forceinline__ uint32_t sel4(bool q0, bool q1, uint32_t A, uint32_t B, uint32_t C, uint32_t D) {
  const uint32_t lo = q0 ? B : A, hi = q0 ? D : C;
  return q1 ? hi : lo;
}

// Single-block Blake2s(m[0..15]) of `t` message bytes (the final block;
// bytes past t are zero in m) by the four lanes of a quad, q = lane & 3; all
// four lanes pass the same 16 message words. Returns output words q and
// 4 + q in lo / hi.
__device__ __forceinline__ void b2_hash_x4(const uint32_t m[16], uint32_t t, uint32_t q, uint32_t& lo, uint32_t& hi) {
  const bool q0 = q & 1, q1 = q & 2;
  const uint32_t iv_lo = sel4(q0, q1, VX_B2_IV0, VX_B2_IV1, VX_B2_IV2, VX_B2_IV3);
  const uint32_t iv_hi = sel4(q0, q1, VX_B2_IV4, VX_B2_IV5, VX_B2_IV6, VX_B2_IV7);
  const uint32_t h_lo = q == 0 ? (VX_B2_IV0 ^ 0x01010020u) : iv_lo, h_hi = iv_hi;
  uint32_t a = h_lo, b = h_hi, c = iv_lo;
  uint32_t d = iv_hi ^ (q == 0 ? t : 0u) ^ (q == 2 ? 0xFFFFFFFFu : 0u);  // v12 ^= t, v14 ^= final
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t x = sel4(q0, q1, m[B2_SIGMA[r][0]], m[B2_SIGMA[r][2]], m[B2_SIGMA[r][4]], m[B2_SIGMA[r][6]]);
    uint32_t y = sel4(q0, q1, m[B2_SIGMA[r][1]], m[B2_SIGMA[r][3]], m[B2_SIGMA[r][5]], m[B2_SIGMA[r][7]]);
    VX_B2_G(a, b, c, d, x, y);
    b = quad_rot(b, VX_DPP_ROT1);
    c = quad_rot(c, VX_DPP_ROT2);
    d = quad_rot(d, VX_DPP_ROT3);
    x = sel4(q0, q1, m[B2_SIGMA[r][8]], m[B2_SIGMA[r][10]], m[B2_SIGMA[r][12]], m[B2_SIGMA[r][14]]);
    y = sel4(q0, q1, m[B2_SIGMA[r][9]], m[B2_SIGMA[r][11]], m[B2_SIGMA[r][13]], m[B2_SIGMA[r][15]]);
    VX_B2_G(a, b, c, d, x, y);
    b = quad_rot(b, VX_DPP_ROT3);
    c = quad_rot(c, VX_DPP_ROT2);
    d = quad_rot(d, VX_DPP_ROT1);
  }
  lo = h_lo ^ a ^ c;
  hi = h_hi ^ b ^ d;
}

// One compression of a chained hash by the four lanes of a quad: lane q
// holds chaining words q and 4 + q in h_lo / h_hi and updates them; t is the
// byte count up to and including this block, last marks the final block.
// All four lanes pass the same 16 message words. Same function as
// b2_compress.
__device__ __forceinline__ void b2_compress_x4(uint32_t& h_lo, uint32_t& h_hi, const uint32_t m[16], uint32_t t,
                                               bool last, uint32_t q) {
  const bool q0 = q & 1, q1 = q & 2;
  uint32_t a = h_lo, b = h_hi;
  uint32_t c = sel4(q0, q1, VX_B2_IV0, VX_B2_IV1, VX_B2_IV2, VX_B2_IV3);
  uint32_t d = sel4(q0, q1, VX_B2_IV4, VX_B2_IV5, VX_B2_IV6, VX_B2_IV7) ^ (q == 0 ? t : 0u) ^
               (q == 2 && last ? 0xFFFFFFFFu : 0u);  // v12 ^= t, v14 ^= final
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t x = sel4(q0, q1, m[B2_SIGMA[r][0]], m[B2_SIGMA[r][2]], m[B2_SIGMA[r][4]], m[B2_SIGMA[r][6]]);
    uint32_t y = sel4(q0, q1, m[B2_SIGMA[r][1]], m[B2_SIGMA[r][3]], m[B2_SIGMA[r][5]], m[B2_SIGMA[r][7]]);
    VX_B2_G(a, b, c, d, x, y);
    b = quad_rot(b, VX_DPP_ROT1);
    c = quad_rot(c, VX_DPP_ROT2);
    d = quad_rot(d, VX_DPP_ROT3);
    x = sel4(q0, q1, m[B2_SIGMA[r][8]], m[B2_SIGMA[r][10]], m[B2_SIGMA[r][12]], m[B2_SIGMA[r][14]]);
    y = sel4(q0, q1, m[B2_SIGMA[r][9]], m[B2_SIGMA[r][11]], m[B2_SIGMA[r][13]], m[B2_SIGMA[r][15]]);
    VX_B2_G(a, b, c, d, x, y);
    b = quad_rot(b, VX_DPP_ROT3);
    c = quad_rot(c, VX_DPP_ROT2);
    d = quad_rot(d, VX_DPP_ROT1);
  }
  h_lo ^= a ^ c;
  h_hi ^= b ^ d;
}

// Merkle node hash (t = 64).
__device__ __forceinline__ void b2_node_x4(const uint32_t m[16], uint32_t q, uint32_t& lo, uint32_t& hi) {
  b2_hash_x4(m, 64u, q, lo, hi);
}

// The 8 output words of a quad hash in every lane of the quad: word k
// (k < 4) is lane k's lo, word 4 + k its hi. The quad must be lanes 0..3 of
// the wavefront (readlane reads fixed lanes).
__device__ __forceinline__ void quad0_gather(uint32_t lo, uint32_t hi, uint32_t out[8]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    out[k] = (uint32_t)__builtin_amdgcn_readlane((int)lo, k);
    out[4 + k] = (uint32_t)__builtin_amdgcn_readlane((int)hi, k);
  }
}

}  // namespace vx
