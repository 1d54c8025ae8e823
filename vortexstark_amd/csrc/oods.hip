// Evaluation of stwo-basis polynomials at a QM31 point, and barycentric sums
// (the stwo plugin's PolyOps::eval_at_point / barycentric_eval_at_point).
//
// Semantics: f(z) = sum_i c_i prod_{bit b of i} F[b] with F[b] =
// factors[log_n - 1 - b]: factors[0] folds the top coefficient bit
// (cuda/circle_ntt.cu:186-231, 429-468; the plugin passes the mappings
// [y, x, pi(x), ...] reversed, crates/cuda-backend/src/poly_ops.rs:382-443).
// All arithmetic is exact, so any summation order gives the same value.
//
// MI355X design: the reference folds level by level (log_n launches, each
// reading and writing a QM31 vector, 16 bytes per element at the first
// level). Here the M31 coefficients are read once. Lane t multiplies its
// 16 consecutive coefficients by the 16 products of F[0..3] (built per
// workgroup in LDS), accumulating 4 products per 64-bit sum before a
// reduction. A 256-lane workgroup then folds its lanes in LDS with F[4..11],
// and the 4096-coefficient block sums are folded 256 at a time with the
// next 8 factors.
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t EV_LANE_LOG = 4;   // coefficients per lane: 16
static constexpr uint32_t EV_WG = 256;       // lanes per workgroup
static constexpr uint32_t EV_WG_LOG = 8;

__device__ __forceinline__ qm31 ld_factor(const uint32_t* factors, uint32_t log_n, uint32_t b) {
  return ld_uniform(reinterpret_cast<const qm31*>(factors) + (log_n - 1 - b));
}

// Folds q[0..W) (W = blockDim.x, a power of two) in LDS with F[fb + j] for
// level j: q[t] += F[fb + j] * q[t + 2^j]. Returns q[0] (valid in lane 0).
__device__ qm31 lds_fold(qm31* q, qm31 v, const uint32_t* factors, uint32_t log_n, uint32_t fb) {
  const uint32_t t = threadIdx.x, W = blockDim.x;
  q[t] = v;
  __syncthreads();
  for (uint32_t j = 0; (1u << j) < W; j++) {
    const qm31 f = ld_factor(factors, log_n, fb + j);
    if ((t & ((2u << j) - 1)) == 0) q[t] = qm_add(q[t], qm_mul(f, q[t + (1u << j)]));
    __syncthreads();
  }
  return q[0];
}

// Block sums of 16 * blockDim.x coefficients each.
__global__ __launch_bounds__(EV_WG) void k_eval_blocks(const uint32_t* __restrict__ c, const uint32_t* factors,
                                                       uint32_t log_n, uint32_t* __restrict__ part) {
  __shared__ qm31 basis[16];
  __shared__ qm31 q[EV_WG];
  const uint32_t t = threadIdx.x;
  for (uint32_t m = t; m < 16; m += blockDim.x) {
    qm31 b{{1u, 0u, 0u, 0u}};
    for (uint32_t k = 0; k < EV_LANE_LOG; k++)
      if ((m >> k) & 1u) b = qm_mul(b, ld_factor(factors, log_n, k));
    basis[m] = b;
  }
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + t;
  const uint4* p = reinterpret_cast<const uint4*>(c + (g << EV_LANE_LOG));
  uint32_t cv[16];
#pragma unroll
  for (uint32_t m = 0; m < 4; m++) {
    const uint4 x = p[m];
    cv[4 * m] = x.x;
    cv[4 * m + 1] = x.y;
    cv[4 * m + 2] = x.z;
    cv[4 * m + 3] = x.w;
  }
  __syncthreads();
  qm31 s;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t grp = 0; grp < 4; grp++) {
      uint64_t acc = 0;  // 4 products < 4 P^2 < 2^64
#pragma unroll
      for (uint32_t m = 0; m < 4; m++) acc += (uint64_t)cv[4 * grp + m] * basis[4 * grp + m].v[k];
      r = m31_add(r, m31_reduce64(acc));
    }
    s.v[k] = r;
  }
  const qm31 v = lds_fold(q, s, factors, log_n, EV_LANE_LOG);
  if (t == 0) reinterpret_cast<qm31*>(part)[blockIdx.x] = v;
}

// Folds blockDim.x consecutive QM31 sums (covering coefficient bits from fb up).
__global__ __launch_bounds__(EV_WG) void k_eval_tree(const uint32_t* __restrict__ in, const uint32_t* factors,
                                                     uint32_t log_n, uint32_t fb, uint32_t* __restrict__ out) {
  __shared__ qm31 q[EV_WG];
  const qm31 v = reinterpret_cast<const qm31*>(in)[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x];
  const qm31 r = lds_fold(q, v, factors, log_n, fb);
  if (threadIdx.x == 0) reinterpret_cast<qm31*>(out)[blockIdx.x] = r;
}

// n < 16: one lane.
__global__ void k_eval_small(const uint32_t* c, const uint32_t* factors, uint32_t log_n, uint32_t* out) {
  qm31 v[8];
  const uint32_t n = 1u << log_n;
  if (log_n == 0) {
    reinterpret_cast<qm31*>(out)[0] = qm31{{c[0], 0u, 0u, 0u}};
    return;
  }
  const qm31 f0 = ld_factor(factors, log_n, 0);
  for (uint32_t i = 0; i < n / 2; i++) v[i] = qm_add(qm31{{c[2 * i], 0u, 0u, 0u}}, qm_scale(f0, c[2 * i + 1]));
  for (uint32_t b = 1, m = n / 2; m > 1; b++, m >>= 1) {
    const qm31 f = ld_factor(factors, log_n, b);
    for (uint32_t i = 0; i < m / 2; i++) v[i] = qm_add(v[2 * i], qm_mul(f, v[2 * i + 1]));
  }
  reinterpret_cast<qm31*>(out)[0] = v[0];
}

const uint32_t* eval_at_point(const uint32_t* c, const uint32_t* factors, uint32_t log_n, uint32_t* s1,
                              uint32_t* s2, hipStream_t s) {
  require(log_n <= 31, "eval_at_point: log_n must be <= 31");
  if (log_n < EV_LANE_LOG) {
    hipLaunchKernelGGL(k_eval_small, dim3(1), dim3(1), 0, s, c, factors, log_n, s1);
    VX_HIP(hipGetLastError());
    return s1;
  }
  const uint32_t lanes_log = log_n - EV_LANE_LOG;
  const uint32_t wg_log = lanes_log < EV_WG_LOG ? lanes_log : EV_WG_LOG;
  uint32_t groups_log = lanes_log - wg_log;
  hipLaunchKernelGGL(k_eval_blocks, dim3(1u << groups_log), dim3(1u << wg_log), 0, s, c, factors, log_n, s1);
  VX_HIP(hipGetLastError());
  uint32_t fb = EV_LANE_LOG + wg_log;
  uint32_t* in = s1;
  uint32_t* out = s2;
  while (groups_log > 0) {
    const uint32_t w = groups_log < EV_WG_LOG ? groups_log : EV_WG_LOG;
    groups_log -= w;
    hipLaunchKernelGGL(k_eval_tree, dim3(1u << groups_log), dim3(1u << w), 0, s, in, factors, log_n, fb, out);
    VX_HIP(hipGetLastError());
    fb += w;
    uint32_t* tmp = in;
    in = out;
    out = tmp;
  }
  return in;
}

// ---------------------------------------------------------------------------
// Barycentric sum (cuda/barycentric_eval.cu:41-99): block b of n_blocks
// 256-lane blocks accumulates evals[i] * weights[i] (QM31 AoS) over the
// grid-stride set i = (b * 256 + t) + k * n_blocks * 256, and writes its
// partial sum out[4b..4b+4). Same partition as the reference, so every
// partial is the same value. Products are accumulated 4 per 64-bit sum.
static constexpr uint32_t BARY_THREADS = 256;

__global__ __launch_bounds__(BARY_THREADS) void k_barycentric(const uint32_t* __restrict__ ev,
                                                              const uint32_t* __restrict__ w, uint32_t n,
                                                              uint32_t* __restrict__ out) {
  __shared__ qm31 q[BARY_THREADS];
  const uint32_t t = threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * BARY_THREADS;
  uint32_t r[4] = {0u, 0u, 0u, 0u};
  uint64_t acc[4] = {0, 0, 0, 0};
  uint32_t pending = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * BARY_THREADS + t; i < n; i += stride) {
    const uint32_t e = ev[i];
    const uint4 x = reinterpret_cast<const uint4*>(w)[i];
    acc[0] += (uint64_t)e * x.x;
    acc[1] += (uint64_t)e * x.y;
    acc[2] += (uint64_t)e * x.z;
    acc[3] += (uint64_t)e * x.w;
    if (++pending == 4) {
      for (int k = 0; k < 4; k++) {
        r[k] = m31_add(r[k], m31_reduce64(acc[k]));
        acc[k] = 0;
      }
      pending = 0;
    }
  }
  for (int k = 0; k < 4; k++) r[k] = m31_add(r[k], m31_reduce64(acc[k]));
  q[t] = qm31{{r[0], r[1], r[2], r[3]}};
  __syncthreads();
  for (uint32_t h = BARY_THREADS / 2; h > 0; h >>= 1) {
    if (t < h) q[t] = qm_add(q[t], q[t + h]);
    __syncthreads();
  }
  if (t == 0) reinterpret_cast<qm31*>(out)[blockIdx.x] = q[0];
}

void barycentric_partials(const uint32_t* ev, const uint32_t* w, uint32_t n, uint32_t* out, uint32_t n_blocks,
                          hipStream_t s) {
  require(n_blocks >= 1 && n_blocks <= 65535, "barycentric: 1..65535 blocks");
  require((reinterpret_cast<uintptr_t>(w) & 15u) == 0, "barycentric: weights must be 16-byte aligned");
  hipLaunchKernelGGL(k_barycentric, dim3(n_blocks), dim3(BARY_THREADS), 0, s, ev, w, n, out);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
