// Fiat-Shamir channel (src/channel.rs:15-185) over word-aligned messages,
// usable from a single device thread and from the host verifier.
//
// The 32-byte state is kept as 8 little-endian words, so every message the
// channel hashes is a sequence of whole words:
//   mix_digest  state || digest                      (64 B, channel.rs:30-38)
//   squeeze     state || n_draws || 0x00             (37 B, channel.rs:177-184)
//   mix_felts   state || felt words ...              (multi-block, :46-59)
//   mix_u64     state || lo || hi                    (40 B, :120-129)
//   PoW prefix  0x12345678 || 0^12 || state || bits  (52 B, :140-168)
#pragma once
#include "vx_blake2s.h"
#include "vx_field.h"

namespace vx {

struct Channel {
  uint32_t state[8];
  uint32_t n_draws;
};

VX_HD void ch_init(Channel& c) {
  for (int i = 0; i < 8; i++) c.state[i] = 0;
  c.n_draws = 0;
}

VX_HD void ch_mix_digest(Channel& c, const uint32_t d[8]) {
  uint32_t m[16];
  for (int i = 0; i < 8; i++) {
    m[i] = c.state[i];
    m[8 + i] = d[i];
  }
  uint32_t h[8];
  b2_init(h);
  b2_compress(h, m, 64u, true);
  for (int i = 0; i < 8; i++) c.state[i] = h[i];
  c.n_draws = 0;
}

VX_HD void ch_squeeze(Channel& c, uint32_t out[8]) {
  uint32_t m[16] = {0};
  for (int i = 0; i < 8; i++) m[i] = c.state[i];
  m[8] = c.n_draws;
  // byte 36 is the 0x00 domain byte; the rest of the block is zero padding
  c.n_draws++;
  b2_init(out);
  b2_compress(out, m, 37u, true);
}

// Rejection-samples until all 8 words are < 2P, then reduces words 0..3.
VX_HD qm31 ch_draw_felt(Channel& c) {
  for (;;) {
    uint32_t w[8];
    ch_squeeze(c, w);
    bool ok = true;
    for (int i = 0; i < 8; i++) ok = ok && (w[i] < 2u * P);
    if (ok) return {{w[0] % P, w[1] % P, w[2] % P, w[3] % P}};
  }
}

VX_HD uint64_t ch_draw_number(Channel& c, uint64_t bound) {
  uint32_t w[8];
  ch_squeeze(c, w);
  uint64_t raw = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  return raw % bound;
}

// mix_felts over n_felts QM31 values given as 4 words each (n_felts <= 8 here;
// the message buffer holds the state plus up to 8 felts = 40 words).
VX_HD void ch_mix_felts(Channel& c, const uint32_t* felt_words, uint32_t n_felts) {
  if (n_felts == 0) return;
  uint32_t msg[8 + 4 * 8];
  for (int i = 0; i < 8; i++) msg[i] = c.state[i];
  for (uint32_t i = 0; i < 4 * n_felts; i++) msg[8 + i] = felt_words[i];
  uint32_t h[8];
  b2_hash_long(msg, 32u + 16u * n_felts, h);
  for (int i = 0; i < 8; i++) c.state[i] = h[i];
  c.n_draws = 0;
}

VX_HD void ch_mix_u64(Channel& c, uint64_t v) {
  uint32_t m[16] = {0};
  for (int i = 0; i < 8; i++) m[i] = c.state[i];
  m[8] = (uint32_t)v;
  m[9] = (uint32_t)(v >> 32);
  uint32_t h[8];
  b2_init(h);
  b2_compress(h, m, 40u, true);
  for (int i = 0; i < 8; i++) c.state[i] = h[i];
  c.n_draws = 0;
}

// PoW prefix digest Blake2s(0x12345678 || 0^12 || state || pow_bits).
VX_HD void ch_pow_prefix(const Channel& c, uint32_t pow_bits, uint32_t out[8]) {
  uint32_t m[16] = {0};
  m[0] = 0x12345678u;
  for (int i = 0; i < 8; i++) m[4 + i] = c.state[i];
  m[12] = pow_bits;
  b2_init(out);
  b2_compress(out, m, 52u, true);
}

// Trailing zeros of the first 16 bytes (u128 LE) of Blake2s(prefix || nonce).
// With M31_OUT every output word is first reduced mod P, as stwo's
// Blake2sChannelGeneric<true> hashes (cuda/grind_m31_output.cu:44-63): a
// 32-bit word x becomes (x & P) + (x >> 31), less P when that reaches P.
template <bool M31_OUT = false>
VX_HD uint32_t pow_trailing_zeros(const uint32_t prefix[8], uint64_t nonce) {
  uint32_t m[16] = {0};
  for (int i = 0; i < 8; i++) m[i] = prefix[i];
  m[8] = (uint32_t)nonce;
  m[9] = (uint32_t)(nonce >> 32);
  uint32_t h[8];
  b2_init(h);
  b2_compress(h, m, 40u, true);
  for (int i = 0; i < 4; i++) {
    uint32_t w = h[i];
    if (M31_OUT) {
      w = (w & P) + (w >> 31);
      w = w >= P ? w - P : w;
    }
    if (w) return 32u * i + (uint32_t)__builtin_ctz(w);
  }
  return 128u;
}

}  // namespace vx
