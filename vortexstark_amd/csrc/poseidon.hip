// Poseidon252 for the stwo plugin's Poseidon252 channel and Merkle
// operations: the PoW grind (cuda/grind_poseidon.cu), lifted leaf hashes and
// parent layers (cuda/merkle_poseidon252.cu), driven by
// crates/cuda-backend/src/grind_ops.rs:263-348 and merkle_ops.rs:230-333.
//
// The field is the Stark prime p = 2^251 + 17 * 2^192 + 1. An element is 8
// little-endian 32-bit words of its Montgomery representative a * 2^256 mod
// p, the same bytes as the plugin's 4 x u64 (starknet-ff's internal form),
// always reduced below p. 32-bit words suit the wavefront's 32 x 32 -> 64
// multiply-add; p's words are 1, 0, 0, 0, 0, 0, 0x11, 0x08000000, so p = 1
// mod 2^32, the Montgomery factor -p^-1 mod 2^32 is -1 and each reduction
// step adds m * p through three nonzero words only.
//
// The permutation is Starknet's Hades (width 3, 4 full + 83 partial + 4
// full rounds, x^3, MDS [[3,1,1],[1,-1,1],[1,1,-2]]) with its uncompressed
// round constants (poseidon_consts.h) held in constant memory and read with
// wave-uniform indices (scalar loads). Every lane owns one whole state: a
// nonce, a leaf or a parent.
#include "vx_internal.h"
#include "poseidon_consts.h"

namespace vx {

namespace {

struct F252 {
  uint32_t w[8];
};

__constant__ F252 c_ark[3 * VX_POSEIDON_ROUNDS] = VX_POSEIDON_ARK_MONT;

constexpr uint32_t PW6 = 0x11u, PW7 = 0x08000000u;  // p's words 6 and 7 (word 0 is 1)
__device__ __forceinline__ constexpr uint32_t pword(int k) { return k == 0 ? 1u : k == 6 ? PW6 : k == 7 ? PW7 : 0u; }

__device__ __forceinline__ F252 f_const(const uint32_t (&v)[8]) {
  F252 r;
#pragma unroll
  for (int k = 0; k < 8; k++) r.w[k] = v[k];
  return r;
}
__device__ __forceinline__ F252 f_one() {
  constexpr uint32_t v[8] = VX_F252_MONT_ONE;
  return f_const(v);
}
__device__ __forceinline__ F252 f_two() {
  constexpr uint32_t v[8] = VX_F252_MONT_TWO;
  return f_const(v);
}
__device__ __forceinline__ F252 f_r2() {
  constexpr uint32_t v[8] = VX_F252_R2;
  return f_const(v);
}

// t in [0, 2p) -> t mod p. Carry and borrow chains use the 32-bit carry
// builtins (v_add_co / v_addc_co, v_sub_co / v_subb_co), no 64-bit temporaries.
// t - p is computed as t + (2^256 - p), whose carry out is set exactly when
// t >= p (an add chain; a borrow chain against p's zero words compiles to
// three instructions per word).
__device__ __forceinline__ void f_reduce_once(F252& t) {
  uint32_t d[8], c = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) d[k] = __builtin_addc(t.w[k], ~pword(k) + (k == 0), c, &c);
#pragma unroll
  for (int k = 0; k < 8; k++) t.w[k] = c ? d[k] : t.w[k];
}

__device__ __forceinline__ F252 f_add(const F252& a, const F252& b) {
  F252 r;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) r.w[k] = __builtin_addc(a.w[k], b.w[k], c, &c);
  f_reduce_once(r);  // a + b < 2p < 2^253: no carry out of word 7
  return r;
}

// Lazy forms for the permutation's linear layer: sums without reduction
// (every operand bound is stated where it is used), k p - x for x < k p, and
// one reduction of a value below 16p to [0, 2p).
__device__ __forceinline__ F252 f_add_nr(const F252& a, const F252& b) {
  F252 r;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) r.w[k] = __builtin_addc(a.w[k], b.w[k], c, &c);
  return r;
}
// 6p = 6 + 102 * 2^192 + 3 * 2^252: words 0, 6 and 7.
__device__ __forceinline__ constexpr uint32_t p6word(int k) { return k == 0 ? 6u : k == 6 ? 0x66u : k == 7 ? 0x30000000u : 0u; }
__device__ __forceinline__ F252 f_6p_minus(const F252& x) {  // x < 6p
  F252 r;
  uint32_t c = 1;  // 6p + ~x + 1
#pragma unroll
  for (int k = 0; k < 8; k++) r.w[k] = __builtin_addc(p6word(k), ~x.w[k], c, &c);
  return r;
}
// x < 16p: with q = x >> 251, x - q p = (x mod 2^251) - q (17 * 2^192 + 1)
// lies in (-2^200, 2^251); p is added back when it is negative.
__device__ __forceinline__ F252 f_red16(const F252& x) {
  const uint32_t q = x.w[7] >> 27;
  F252 r;
  uint32_t c = 1;  // low + ~(q + 17 q 2^192) + 1; carry out = no borrow
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t xk = k == 7 ? (x.w[7] & 0x07FFFFFFu) : x.w[k];
    const uint32_t sk = k == 0 ? q : (k == 6 ? 17u * q : 0u);
    r.w[k] = __builtin_addc(xk, ~sk, c, &c);
  }
  const uint32_t mask = c - 1u;  // all ones on a borrow
  uint32_t d = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) r.w[k] = __builtin_addc(r.w[k], pword(k) & mask, d, &d);
  return r;
}

// Montgomery product a * b / 2^256 mod p, product scanning: column k of
// a * b + m * p accumulates in a 64-bit word plus an overflow word (each
// term one multiply-add into the accumulator and one add of its carry), so
// no partial product is split or zero-extended. m_k = -(column k) mod 2^32
// clears the column (p = 1 mod 2^32); m * p adds only at p's words 0, 6, 7.
// The result is below 2p, then reduced once.
__device__ __forceinline__ void mac(uint64_t& acc, uint32_t& ovf, uint32_t x, uint32_t y) {
  const uint64_t prev = acc;
  acc = (uint64_t)x * y + acc;  // one v_mad_u64_u32 into the accumulator
  ovf += acc < prev;
}
// Inputs below 4p give a result below 2p (a b / 2^256 < 2^250 < p/2, plus
// m p / 2^256 < p): the permutation keeps its states in [0, 2p) and leaves
// this product unreduced.
__device__ __forceinline__ F252 f_mul_lazy(const F252& a, const F252& b) {
  uint32_t m[8], r[8];
  uint64_t acc = 0;
  uint32_t ovf = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (k - j >= 0 && k - j < 8) mac(acc, ovf, a.w[j], b.w[k - j]);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (k - j == 6 && j < k) mac(acc, ovf, m[j], PW6);
      if (k - j == 7 && j < k) mac(acc, ovf, m[j], PW7);
    }
    if (k < 8) {
      m[k] = 0u - (uint32_t)acc;
      unsigned long long c;
      acc = __builtin_addcll(acc, (uint64_t)m[k], 0, &c);
      ovf += (uint32_t)c;
    } else {
      r[k - 8] = (uint32_t)acc;
    }
    acc = (acc >> 32) | ((uint64_t)ovf << 32);
    ovf = 0;
  }
  F252 t;
#pragma unroll
  for (int k = 0; k < 8; k++) t.w[k] = r[k];
  return t;
}
__device__ __forceinline__ F252 f_mul(const F252& a, const F252& b) {
  F252 t = f_mul_lazy(a, b);
  f_reduce_once(t);
  return t;
}

// x below 4p -> x^3 below 2p.
__device__ __forceinline__ F252 f_cube(const F252& x) { return f_mul_lazy(f_mul_lazy(x, x), x); }

// s <- M s with M = [[3,1,1],[1,-1,1],[1,1,-2]] for s0, s1 < 3p and s2 < 2p
// (s2 is always an S-box output): t < 8p, and each row, with 6p added
// before a subtraction, stays below 14p before its one reduction to [0, 2p).
__device__ __forceinline__ void mix(F252 (&s)[3]) {
  const F252 t = f_add_nr(f_add_nr(s[0], s[1]), s[2]);
  const F252 s0 = f_red16(f_add_nr(t, f_add_nr(s[0], s[0])));
  const F252 s1 = f_red16(f_add_nr(t, f_6p_minus(f_add_nr(s[1], s[1]))));
  const F252 s2 = f_red16(f_add_nr(t, f_6p_minus(f_add_nr(f_add_nr(s[2], s[2]), s[2]))));
  s[0] = s0;
  s[1] = s1;
  s[2] = s2;
}

// States in [0, 2p); a round constant (< p) brings them below 3p.
__device__ __forceinline__ void full_round(F252 (&s)[3], uint32_t r) {
#pragma unroll
  for (int j = 0; j < 3; j++) s[j] = f_cube(f_add_nr(s[j], c_ark[3 * r + j]));
  mix(s);
}

__device__ __forceinline__ void partial_round(F252 (&s)[3], uint32_t r) {
  s[0] = f_add_nr(s[0], c_ark[3 * r]);
  s[1] = f_add_nr(s[1], c_ark[3 * r + 1]);
  s[2] = f_cube(f_add_nr(s[2], c_ark[3 * r + 2]));
  mix(s);
}

// Canonical states in and out; between the rounds they stay in [0, 2p).
// The loops stay rolled: one round's code is a few thousand instructions.
__device__ __forceinline__ void permute(F252 (&s)[3]) {
#pragma unroll 1
  for (uint32_t r = 0; r < VX_POSEIDON_FULL_HALF; r++) full_round(s, r);
#pragma unroll 1
  for (uint32_t r = VX_POSEIDON_FULL_HALF; r < VX_POSEIDON_FULL_HALF + VX_POSEIDON_PARTIAL; r++) partial_round(s, r);
#pragma unroll 1
  for (uint32_t r = VX_POSEIDON_FULL_HALF + VX_POSEIDON_PARTIAL; r < VX_POSEIDON_ROUNDS; r++) full_round(s, r);
#pragma unroll
  for (int j = 0; j < 3; j++) f_reduce_once(s[j]);
}

// Elements are 4 x u64 (8-byte aligned): four 8-byte accesses each.
__device__ __forceinline__ F252 f_load(const uint64_t* p) {
  F252 r;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint64_t v = p[k];
    r.w[2 * k] = (uint32_t)v;
    r.w[2 * k + 1] = (uint32_t)(v >> 32);
  }
  return r;
}
__device__ __forceinline__ void f_store(uint64_t* p, const F252& x) {
#pragma unroll
  for (int k = 0; k < 4; k++) p[k] = (uint64_t)x.w[2 * k] | ((uint64_t)x.w[2 * k + 1] << 32);
}

// ---------------------------------------------------------------- grind
// poseidon_hash(prefixed, nonce) = permute([prefixed, nonce, 2])[0]; the
// test counts the trailing zeros of its canonical value's low 128 bits
// (cuda/grind_poseidon.cu:1-21). Nonces are scanned in increasing order with
// a wave-uniform exit once the wave's next nonces pass the best hit, so the
// result is the smallest valid nonce of the range.
constexpr uint64_t POSEIDON_GRIND_LIMIT = 1ull << 44;  // exit bound every lane reaches

__device__ __forceinline__ uint32_t pow_zeros(const F252& prefix, uint64_t nonce) {
  F252 s[3];
  const F252 n = {{(uint32_t)nonce, (uint32_t)(nonce >> 32), 0, 0, 0, 0, 0, 0}};
  s[0] = prefix;
  s[1] = f_mul(n, f_r2());  // to Montgomery form
  s[2] = f_two();
  permute(s);
  const F252 one = {{1, 0, 0, 0, 0, 0, 0, 0}};
  const F252 h = f_mul(s[0], one);  // out of Montgomery form
  uint32_t z = 0;
  bool done = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!done) {
      if (h.w[k]) {
        z += __builtin_ctz(h.w[k]);
        done = true;
      } else {
        z += 32;
      }
    }
  }
  return z;
}

__device__ __forceinline__ uint64_t wave_min(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint64_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_poseidon_grind(const uint64_t* prefix_mont, uint32_t bits,
                                                        unsigned long long* best, uint64_t start, uint64_t count) {
  const F252 prefix = f_load(prefix_mont);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t end = start + count;
  for (uint64_t base = start + (uint64_t)blockIdx.x * blockDim.x; base < end; base += stride) {
    const uint64_t seen = __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t cur_best = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(seen >> 32)) << 32) |
                              (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)seen);
    if (base > cur_best) break;
    const uint64_t n = base + threadIdx.x;
    uint64_t hit = ~0ull;
    if (n < end && pow_zeros(prefix, n) >= bits) hit = n;
    hit = wave_min(hit);
    if ((threadIdx.x & 63) == 0 && hit != ~0ull) atomicMin(best, (unsigned long long)hit);
  }
}

__global__ void k_poseidon_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------- Merkle
// Lifted row of a column of 2^col_log rows under a tree of 2^lifting leaves
// (merkle_poseidon252.cu:28-33); masked so a one-row column reads row 0.
__device__ __forceinline__ uint32_t lifted_row252(uint32_t leaf, uint32_t lifting, uint32_t col_log) {
  if (col_log >= lifting) return leaf;
  const uint32_t r = ((leaf >> (lifting - col_log + 1)) << 1) | (leaf & 1);
  return r & ((1u << col_log) - 1);
}

// n (1..8, wave-uniform) M31 values, first most significant, 31 bits each;
// fewer than 8 add n * 2^248. Below 2^251 < p, then to Montgomery form.
__device__ __forceinline__ F252 pack_m31s(const uint32_t (&v)[8], uint32_t n) {
  F252 x = {{0, 0, 0, 0, 0, 0, 0, 0}};
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if ((uint32_t)k < n) {
#pragma unroll
      for (int j = 7; j > 0; j--) x.w[j] = (x.w[j] << 31) | (x.w[j - 1] >> 1);
      x.w[0] = (x.w[0] << 31) | v[k];
    }
  }
  if (n < 8) x.w[7] += n << 24;  // n * 2^248
  return f_mul(x, f_r2());
}

// Leaf hash: poseidon_hash_many of the packed values of every column at the
// leaf's lifted row, 8 columns per element (merkle_poseidon252.cu:37-123).
__global__ __launch_bounds__(256) void k_poseidon_leaves(const uint32_t* const* cols, const uint32_t* logs,
                                                         uint32_t n_cols, uint32_t lifting, uint64_t* out,
                                                         uint32_t n_leaves) {
  const uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= n_leaves) return;
  F252 s[3];
#pragma unroll
  for (int j = 0; j < 3; j++) s[j] = {{0, 0, 0, 0, 0, 0, 0, 0}};
  const uint32_t n_felts = (n_cols + 7) / 8;
  F252 first = s[0];
  for (uint32_t f = 0; f < n_felts; f++) {
    const uint32_t c0 = 8 * f, n = n_cols - c0 < 8 ? n_cols - c0 : 8;
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      v[k] = 0;
      if ((uint32_t)k < n) v[k] = cols[c0 + k][lifted_row252(leaf, lifting, logs[c0 + k])];
    }
    const F252 fe = pack_m31s(v, n);
    if ((f & 1) == 0) {
      first = fe;
    } else {
      s[0] = f_add(s[0], first);
      s[1] = f_add(s[1], fe);
      permute(s);
    }
  }
  if (n_felts & 1) {  // a lone last element pads the second rate word
    s[0] = f_add(s[0], first);
    s[1] = f_add(s[1], f_one());
  } else {
    s[0] = f_add(s[0], f_one());
  }
  permute(s);
  f_store(out + 4 * (uint64_t)leaf, s[0]);
}

// Parent i = poseidon_hash(prev[2i], prev[2i + 1]) (merkle_poseidon252.cu:127-148).
__global__ __launch_bounds__(256) void k_poseidon_next_layer(const uint64_t* prev, uint64_t* out, uint32_t n_parents) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_parents) return;
  F252 s[3] = {f_load(prev + 8 * (uint64_t)i), f_load(prev + 8 * (uint64_t)i + 4), f_two()};
  permute(s);
  f_store(out + 4 * (uint64_t)i, s[0]);
}

uint32_t poseidon_grind_grid(uint32_t bits) {
  // a round of the grid tests grid x 256 nonces; at about 10^5 instructions
  // per nonce the whole GPU is used from PoW 18 up
  const uint64_t wgs = bits >= 18 ? 2048ull : ((1ull << bits) / 4 / 256);
  return (uint32_t)(wgs < 64 ? 64 : (wgs > 2048 ? 2048 : wgs));
}

}  // namespace

void poseidon_grind_batch(const uint64_t* d_prefix_mont, uint64_t* d_result, uint32_t pow_bits, uint64_t batch_offset,
                          uint64_t n, hipStream_t s) {
  require(d_prefix_mont && d_result && n > 0, "grind_pow_poseidon: bad arguments");
  require(pow_bits <= 128, "grind_pow_poseidon: pow_bits must be <= 128");
  require(batch_offset + n >= batch_offset, "grind_pow_poseidon: nonce range overflows");
  const uint64_t blocks = (n + 255) / 256;
  const uint32_t grid = (uint32_t)(blocks < 2048 ? blocks : 2048);
  hipLaunchKernelGGL(k_poseidon_grind, dim3(grid), dim3(256), 0, s, d_prefix_mont, pow_bits,
                     reinterpret_cast<unsigned long long*>(d_result), batch_offset, n);
  VX_HIP(hipGetLastError());
}

void poseidon_grind_min(const uint64_t* d_prefix_mont, uint32_t pow_bits, uint64_t* d_result, hipStream_t s) {
  require(d_prefix_mont && d_result, "grind_pow_stwo_poseidon: null pointer");
  require(pow_bits >= 1 && pow_bits <= 40, "grind_pow_stwo_poseidon: pow_bits must be 1..40");
  auto* best = reinterpret_cast<unsigned long long*>(d_result);
  hipLaunchKernelGGL(k_poseidon_set_u64, dim3(1), dim3(64), 0, s, best, ~0ull);
  hipLaunchKernelGGL(k_poseidon_grind, dim3(poseidon_grind_grid(pow_bits)), dim3(256), 0, s, d_prefix_mont, pow_bits,
                     best, 0ull, POSEIDON_GRIND_LIMIT);
  VX_HIP(hipGetLastError());
}

void poseidon_leaves(const uint32_t* const* d_cols, const uint32_t* d_logs, uint32_t n_cols, uint32_t lifting_log,
                     uint64_t* d_out, uint32_t n_leaves, hipStream_t s) {
  require(d_out && (n_cols == 0 || (d_cols && d_logs)), "build_leaves_poseidon252: null pointer");
  require(lifting_log < 32 && (uint64_t)n_leaves <= (1ull << lifting_log),
          "build_leaves_poseidon252: n_leaves must be <= 2^lifting_log_size < 2^32");
  if (n_leaves == 0) return;
  hipLaunchKernelGGL(k_poseidon_leaves, dim3((n_leaves + 255) / 256), dim3(256), 0, s, d_cols, d_logs, n_cols,
                     lifting_log, d_out, n_leaves);
  VX_HIP(hipGetLastError());
}

void poseidon_next_layer(const uint64_t* d_prev, uint64_t* d_out, uint32_t n_parents, hipStream_t s) {
  require(d_prev && d_out, "build_next_layer_poseidon252: null pointer");
  if (n_parents == 0) return;
  hipLaunchKernelGGL(k_poseidon_next_layer, dim3((n_parents + 255) / 256), dim3(256), 0, s, d_prev, d_out, n_parents);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
