// Blake2s-256 compression for the gfx950 kernels and the host.
//
// Plain unkeyed Blake2s (RFC 7693) as the reference uses it for Merkle
// leaves/nodes and the Fiat-Shamir channel (cuda/include/blake2s.cuh:53-136,
// src/channel.rs:216-298): h0 = IV0 ^ 0x01010020, t = message byte count,
// final flag on the last block. Messages here are always whole 32-bit words.
//
// On CDNA4 the whole state lives in VGPRs; rotations lower to v_alignbit_b32
// and a + b + m to v_add3_u32. The round loop is fully unrolled so every
// message-schedule index is a compile-time register selection.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef VX_HD
#define VX_HD __host__ __device__ __forceinline__
#endif

namespace vx {

#define VX_B2_IV0 0x6A09E667u
#define VX_B2_IV1 0xBB67AE85u
#define VX_B2_IV2 0x3C6EF372u
#define VX_B2_IV3 0xA54FF53Au
#define VX_B2_IV4 0x510E527Fu
#define VX_B2_IV5 0x9B05688Cu
#define VX_B2_IV6 0x1F83D9ABu
#define VX_B2_IV7 0x5BE0CD19u
// h[0] of an unkeyed 32-byte-digest Blake2s: IV0 ^ 0x01010020.
#define VX_B2_H0 (VX_B2_IV0 ^ 0x01010020u)

VX_HD uint32_t rotr32(uint32_t x, uint32_t n) { return (x >> n) | (x << (32u - n)); }

#define VX_B2_G(a, b, c, d, x, y) \
  do {                            \
    a = a + b + (x);              \
    d = rotr32(d ^ a, 16);        \
    c = c + d;                    \
    b = rotr32(b ^ c, 12);        \
    a = a + b + (y);              \
    d = rotr32(d ^ a, 8);         \
    c = c + d;                    \
    b = rotr32(b ^ c, 7);         \
  } while (0)

// One round with message-schedule row s0..s15 (RFC 7693 sigma).
#define VX_B2_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  do {                                                                                   \
    VX_B2_G(v0, v4, v8, v12, m[s0], m[s1]);                                              \
    VX_B2_G(v1, v5, v9, v13, m[s2], m[s3]);                                              \
    VX_B2_G(v2, v6, v10, v14, m[s4], m[s5]);                                             \
    VX_B2_G(v3, v7, v11, v15, m[s6], m[s7]);                                             \
    VX_B2_G(v0, v5, v10, v15, m[s8], m[s9]);                                             \
    VX_B2_G(v1, v6, v11, v12, m[s10], m[s11]);                                           \
    VX_B2_G(v2, v7, v8, v13, m[s12], m[s13]);                                            \
    VX_B2_G(v3, v4, v9, v14, m[s14], m[s15]);                                            \
  } while (0)

// h <- compress(h, m, t, last). `m` is 16 words; `t` is the byte counter.
VX_HD void b2_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, bool last) {
  uint32_t v0 = h[0], v1 = h[1], v2 = h[2], v3 = h[3], v4 = h[4], v5 = h[5], v6 = h[6], v7 = h[7];
  uint32_t v8 = VX_B2_IV0, v9 = VX_B2_IV1, v10 = VX_B2_IV2, v11 = VX_B2_IV3;
  uint32_t v12 = VX_B2_IV4 ^ t, v13 = VX_B2_IV5, v14 = last ? ~VX_B2_IV6 : VX_B2_IV6, v15 = VX_B2_IV7;
  VX_B2_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  VX_B2_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3);
  VX_B2_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4);
  VX_B2_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8);
  VX_B2_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13);
  VX_B2_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9);
  VX_B2_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11);
  VX_B2_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10);
  VX_B2_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5);
  VX_B2_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0);
  h[0] ^= v0 ^ v8;
  h[1] ^= v1 ^ v9;
  h[2] ^= v2 ^ v10;
  h[3] ^= v3 ^ v11;
  h[4] ^= v4 ^ v12;
  h[5] ^= v5 ^ v13;
  h[6] ^= v6 ^ v14;
  h[7] ^= v7 ^ v15;
}

VX_HD void b2_init(uint32_t h[8]) {
  h[0] = VX_B2_H0;
  h[1] = VX_B2_IV1;
  h[2] = VX_B2_IV2;
  h[3] = VX_B2_IV3;
  h[4] = VX_B2_IV4;
  h[5] = VX_B2_IV5;
  h[6] = VX_B2_IV6;
  h[7] = VX_B2_IV7;
}

// Hash of a single-block message of `nwords` (<= 16) words: a Merkle leaf of
// nwords M31 values (cuda/blake2s.cu:13-61; t = 4 * nwords).
VX_HD void b2_hash_words(const uint32_t* w, uint32_t nwords, uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; i++) m[i] = (uint32_t)i < nwords ? w[i] : 0u;
  b2_init(out);
  b2_compress(out, m, 4u * nwords, true);
}

// Merkle node: Blake2s(left32 || right32) (cuda/blake2s.cu:64-87; t = 64).
VX_HD void b2_hash_node(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    m[i] = l[i];
    m[8 + i] = r[i];
  }
  b2_init(out);
  b2_compress(out, m, 64u, true);
}

// Leaf of one M31 value (t = 4) and of one QM31 value (t = 16).
VX_HD void b2_leaf1(uint32_t v, uint32_t out[8]) {
  uint32_t m[16] = {v, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  b2_init(out);
  b2_compress(out, m, 4u, true);
}
VX_HD void b2_leaf4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t out[8]) {
  uint32_t m[16] = {a, b, c, d, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  b2_init(out);
  b2_compress(out, m, 16u, true);
}

// Multi-block Blake2s over a word-aligned message of `nbytes` bytes, where
// nbytes is a multiple of 4 (Channel::mix_felts, src/channel.rs:46-59).
VX_HD void b2_hash_long(const uint32_t* w, uint32_t nbytes, uint32_t out[8]) {
  b2_init(out);
  uint32_t t = 0, off = 0;
  uint32_t m[16];
  while (nbytes - off > 64) {
    for (int i = 0; i < 16; i++) m[i] = w[off / 4 + i];
    t += 64;
    b2_compress(out, m, t, false);
    off += 64;
  }
  uint32_t rem = nbytes - off;
  for (uint32_t i = 0; i < 16; i++) m[i] = 4 * i < rem ? w[off / 4 + i] : 0u;
  t += rem;
  b2_compress(out, m, t, true);
}

}  // namespace vx
