// Fiat-Shamir steps of vx_channel.h computed by the quad of lanes 0..3 of a
// wavefront (device only): each compression is a quad hash (b2_hash_x4) and
// the 8 output words are gathered back into every lane of the quad, so all
// four lanes pass and receive the same channel. Same results as the
// single-lane ch_mix_digest / ch_draw_felt (src/channel.rs:30-38,177-184).
#pragma once
#include "vx_blake2s_x4.h"
#include "vx_channel.h"

namespace vx {

__device__ __forceinline__ void ch_mix_digest_q0(Channel& c, const uint32_t d[8], uint32_t q) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    m[i] = c.state[i];
    m[8 + i] = d[i];
  }
  uint32_t lo, hi;
  b2_hash_x4(m, 64u, q, lo, hi);
  quad0_gather(lo, hi, c.state);
  c.n_draws = 0;
}

// Rejection sampling as ch_draw_felt; the test runs on gathered words, so
// the four lanes leave the loop together.
__device__ __forceinline__ qm31 ch_draw_felt_q0(Channel& c, uint32_t q) {
  for (;;) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = c.state[i];
    m[8] = c.n_draws;
#pragma unroll
    for (int i = 9; i < 16; i++) m[i] = 0;
    c.n_draws++;
    uint32_t lo, hi, w[8];
    b2_hash_x4(m, 37u, q, lo, hi);
    quad0_gather(lo, hi, w);
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 8; i++) ok = ok && (w[i] < 2u * P);
    if (ok) return {{w[0] % P, w[1] % P, w[2] % P, w[3] % P}};
  }
}

}  // namespace vx
