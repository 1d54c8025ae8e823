// Device-resident Fiat-Shamir transcript and deterministic PoW grind.
//
// The transcript (src/channel.rs:15-185) runs on the GPU so a whole proof is
// one stream of kernels with no host round trip: each commitment root is
// mixed by a single lane, and the next fold reads its alpha from device
// memory. Query draws are independent squeezes (n_draws = 0..nq-1 after
// mix_felts), so one lane per query computes them in parallel.
#include "vx_internal.h"

namespace vx {

__global__ void k_channel_step(Channel* ch, int init, const uint32_t* digest, uint32_t* alpha_out,
                               uint32_t* alpha_sq_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Channel c;
  if (init) ch_init(c);
  else c = *ch;
  if (digest) {
    uint32_t d[8];
    for (int k = 0; k < 8; k++) d[k] = digest[k];
    ch_mix_digest(c, d);
  }
  if (alpha_out) {
    qm31 a = ch_draw_felt(c);
    for (int k = 0; k < 4; k++) alpha_out[k] = a.v[k];
    if (alpha_sq_out) {
      qm31 a2 = qm_mul(a, a);
      for (int k = 0; k < 4; k++) alpha_sq_out[k] = a2.v[k];
    }
  }
  *ch = c;
}

void channel_step(Channel* ch, bool init, const uint32_t* digest, uint32_t* alpha_out, uint32_t* alpha_sq_out,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_channel_step, dim3(1), dim3(64), 0, s, ch, init ? 1 : 0, digest, alpha_out, alpha_sq_out);
  VX_HIP(hipGetLastError());
}

// mix_felts(last layer) (src/prover.rs:554); writes n_last + the 8 values
// into the proof and, for PoW, the prefix digest of the new state.
__global__ void k_mix_last(QueryArgs q, uint32_t* prefix_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Channel c = *q.ch;
  uint32_t w[32];
  for (int i = 0; i < 8; i++)
    for (int k = 0; k < 4; k++) w[4 * i + k] = q.last[k][i];
  ch_mix_felts(c, w, 8);
  q.proof[q.last_word] = 8;
  for (int i = 0; i < 32; i++) q.proof[q.last_word + 1 + i] = w[i];
  if (q.pow_bits && prefix_out) ch_pow_prefix(c, q.pow_bits, prefix_out);
  *q.ch = c;
}
void channel_mix_last(const QueryArgs& q, uint32_t* prefix_out, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_last, dim3(1), dim3(64), 0, s, q, prefix_out);
  VX_HIP(hipGetLastError());
}

// Query i is draw_number(domain) at n_draws = i (src/prover.rs:555-557),
// after mix_u64(nonce) when PoW is on.
__global__ void k_draw_queries(QueryArgs q) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  Channel c = *q.ch;
  if (q.pow_bits) {
    uint64_t nonce = *q.nonce;
    ch_mix_u64(c, nonce);
    if (i == 0) {
      q.proof[q.nonce_word] = (uint32_t)nonce;
      q.proof[q.nonce_word + 1] = (uint32_t)(nonce >> 32);
    }
  }
  if (i == 0) q.proof[q.query_word] = q.nq;
  if (i >= q.nq) return;
  c.n_draws = i;
  uint32_t v = (uint32_t)ch_draw_number(c, q.domain);
  q.queries[i] = v;
  q.proof[q.query_word + 1 + i] = v;
}
void channel_draw_queries(const QueryArgs& q, hipStream_t s) {
  require(q.nq >= 1 && q.nq <= 65536, "queries: 1..65536 queries");
  uint32_t blocks = (q.nq + 255) / 256;
  hipLaunchKernelGGL(k_draw_queries, dim3(blocks), dim3(256), 0, s, q);
  VX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ PoW
// Nonces are scanned in increasing order by a grid-stride loop; every lane
// keeps going while its next nonce is below the best found so far, so every
// nonce smaller than the final answer has been tested: the result is the
// smallest valid nonce, independent of scheduling. Each wave reduces its
// hits with a wavefront min before one atomicMin.
static constexpr uint64_t GRIND_LIMIT = 1ull << 44;  // exit bound every lane reaches

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// The stwo plugin's nonce family (crates/cuda-backend/src/grind_ops.rs:
// 71-128): index i stands for nonce ((i >> 20) << 32) | (i & (2^20 - 1)), a
// map that keeps the order, so the smallest index is the smallest nonce.
__device__ __forceinline__ uint64_t stwo_nonce(uint64_t i) { return ((i >> 20) << 32) | (i & 0xFFFFFull); }

template <bool STWO, bool M31_OUT = false>
__global__ __launch_bounds__(256) void k_grind(const uint32_t* prefix_dev, uint32_t bits,
                                              unsigned long long* best, uint64_t start, uint64_t count) {
  uint32_t prefix[8];
  for (int k = 0; k < 8; k++) prefix[k] = prefix_dev[k];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t end = start + count;
  for (uint64_t base = start + (uint64_t)blockIdx.x * blockDim.x; base < end; base += stride) {
    // Wave-uniform exit: lane 0's view of the best hit decides for the wave,
    // so the shuffles below always run with all 64 lanes active.
    const uint64_t seen = __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t cur_best = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(seen >> 32)) << 32) |
                              (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)seen);
    if (base > cur_best) break;
    const uint64_t n = base + threadIdx.x;
    uint64_t hit = ~0ull;
    if (n < end && pow_trailing_zeros<M31_OUT>(prefix, STWO ? stwo_nonce(n) : n) >= bits) hit = n;
    hit = wave_min_u64(hit);
    if ((threadIdx.x & 63) == 0 && hit != ~0ull) atomicMin(best, (unsigned long long)hit);
  }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
__global__ void k_stwo_index_to_nonce(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && *p != ~0ull) *p = stwo_nonce(*p);
}

// Workgroups of a search for `bits` trailing zeros. Every lane of the grid
// tests one nonce per round until the round passes the smallest hit, so a
// search tests about (hit + one round of the grid) nonces; a round of a
// quarter of the expected 2^bits nonces keeps that overshoot small (PoW 20:
// 1024 workgroups, 2^18 lanes), and large searches use the whole GPU
// (2048 workgroups, 8 per CU).
static uint32_t grind_grid(uint32_t bits) {
  const uint64_t wgs = bits >= 20 + 1 ? 2048ull : (1ull << bits) / 4 / 256;
  return (uint32_t)(wgs < 64 ? 64 : (wgs > 2048 ? 2048 : wgs));
}

// Whole search over the stwo family: the same increasing-index scan, then
// the winning index is turned into its nonce. Indices up to 2^44 cover
// hi < 2^24.
// m31_out selects the Blake2sChannelGeneric<true> test (output words
// reduced mod P before the trailing-zero count).
void grind_pow_stwo(const uint32_t* d_prefix, uint32_t pow_bits, uint64_t* d_result, bool m31_out, hipStream_t s) {
  require(pow_bits >= 1 && pow_bits <= 32, "grind_pow_stwo: pow_bits must be 1..32");
  auto* best = reinterpret_cast<unsigned long long*>(d_result);
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, best, ~0ull);
  if (m31_out)
    hipLaunchKernelGGL((k_grind<true, true>), dim3(grind_grid(pow_bits)), dim3(256), 0, s, d_prefix, pow_bits, best,
                       0ull, GRIND_LIMIT);
  else
    hipLaunchKernelGGL((k_grind<true, false>), dim3(grind_grid(pow_bits)), dim3(256), 0, s, d_prefix, pow_bits,
                       best, 0ull, GRIND_LIMIT);
  hipLaunchKernelGGL(k_stwo_index_to_nonce, dim3(1), dim3(64), 0, s, best);
  VX_HIP(hipGetLastError());
}

void grind_pow(const uint32_t* d_prefix, uint32_t pow_bits, uint64_t* d_result, hipStream_t s) {
  require(pow_bits >= 1 && pow_bits <= 32, "grind_pow: pow_bits must be 1..32");
  auto* best = reinterpret_cast<unsigned long long*>(d_result);
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, best, ~0ull);
  // every lane loops until its nonce exceeds the best hit
  hipLaunchKernelGGL(k_grind<false>, dim3(grind_grid(pow_bits)), dim3(256), 0, s, d_prefix, pow_bits, best, 0ull,
                     GRIND_LIMIT);
  VX_HIP(hipGetLastError());
}

// Reference-style batch (cuda/grind.cu:94-110): test nonces
// [batch_offset, batch_offset + n) and atomicMin hits into *d_result.
void grind_pow_batch(const uint32_t* d_prefix, uint64_t* d_result, uint32_t pow_bits, uint64_t batch_offset,
                     uint64_t n, bool m31_out, hipStream_t s) {
  require(pow_bits <= 128 && n > 0, "grind_pow_batch: bad arguments");
  uint64_t blocks = (n + 255) / 256;
  uint32_t grid = (uint32_t)(blocks < 4096 ? blocks : 4096);
  auto* best = reinterpret_cast<unsigned long long*>(d_result);
  if (m31_out)
    hipLaunchKernelGGL((k_grind<false, true>), dim3(grid), dim3(256), 0, s, d_prefix, pow_bits, best, batch_offset,
                       n);
  else
    hipLaunchKernelGGL((k_grind<false, false>), dim3(grid), dim3(256), 0, s, d_prefix, pow_bits, best, batch_offset,
                       n);
  VX_HIP(hipGetLastError());
}

struct WordWrites {
  uint32_t n;
  uint32_t off[16];
  uint32_t val[16];
};
__global__ void k_write_words(uint32_t* dst, WordWrites w) {
  uint32_t i = threadIdx.x;
  if (i < w.n) dst[w.off[i]] = w.val[i];
}
void write_header(uint32_t* proof, const uint32_t* words, uint32_t n, hipStream_t s) {
  // words: n (offset, value) pairs
  require(n <= 16, "write_header: at most 16 words");
  WordWrites w{};
  w.n = n;
  for (uint32_t i = 0; i < n; i++) {
    w.off[i] = words[2 * i];
    w.val[i] = words[2 * i + 1];
  }
  hipLaunchKernelGGL(k_write_words, dim3(1), dim3(64), 0, s, proof, w);
  VX_HIP(hipGetLastError());
}

// Proof download by the GPU itself: the workgroups store the proof
// straight into pinned host memory (mapped into the device's address space),
// 16 bytes per lane per step, instead of a DMA copy whose start-up cost
// dominated for 1-2 MB proofs.
__global__ __launch_bounds__(256) void k_copy_to_host(uint32_t* dst, const uint32_t* src, uint64_t words) {
  const uint64_t n16 = words >> 2;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint4* d4 = reinterpret_cast<uint4*>(dst);
  const uint4* s4 = reinterpret_cast<const uint4*>(src);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) d4[i] = s4[i];
  const uint64_t tail = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tail < (words & 3)) dst[4 * n16 + tail] = src[4 * n16 + tail];
}

void copy_to_host(uint32_t* host_dst, const uint32_t* dev_src, uint64_t words, hipStream_t s) {
  require((reinterpret_cast<uintptr_t>(host_dst) & 15u) == 0 && (reinterpret_cast<uintptr_t>(dev_src) & 15u) == 0,
          "copy_to_host: 16-byte aligned buffers");
  const uint64_t blocks = ((words >> 2) + 255) / 256;
  const uint32_t grid = (uint32_t)(blocks < 1 ? 1 : (blocks > 512 ? 512 : blocks));
  hipLaunchKernelGGL(k_copy_to_host, dim3(grid), dim3(256), 0, s, host_dst, dev_src, words);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
