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
// level). Here the M31 coefficients are read once, by k_eval_big (n >= 2^14:
// 64 coefficients per lane, coalesced loads, the basis as scalars) or
// k_eval_blocks (smaller n: 16 consecutive coefficients per lane against the
// 16 products of F[0..3] in LDS). Products are accumulated 4 per 64-bit sum
// before a reduction; a workgroup folds its lanes in LDS, and the block sums
// are folded 256 at a time with the next 8 factors.
#include "prover.h"
#include "vx_internal.h"
#include "vx_blake2s_x4.h"

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

// ---- n >= 2^14: 64 coefficients per lane, basis in scalar registers.
// A workgroup covers 2^14 coefficients. Lane t reads the 16 uint4 at
// t + 256 m (m < 16), fully coalesced, so coefficient r + 4t + 1024m sits in
// component r of load m: bits 0-1 = r, 2-9 = t, 10-13 = m. The setup kernel
// tabulates E[r + 4 m01] = prod of F[0], F[1], F[10], F[11] over the set
// bits of (r, m01) and D[m23] over F[12], F[13]; lanes read them as
// wave-uniform scalars. Per lane: four 16-term sums w[m23] against E, then
// w0 + D1 w1 + D2 w2 + D3 w3; the 256 lanes are folded with F[2..9].
// (Per 2^14 coefficients a wave issues about 1600 VALU instructions, a
// quarter of the 16-per-lane kernel's; at 2^24 this kernel takes 27 us
// against 50 us. Looping each workgroup over several chunks with the next
// chunk's loads in flight measured the same 27 us.)
static constexpr uint32_t EV_BIG_LOG = 14;

__global__ void k_eval_setup(const uint32_t* factors, uint32_t log_n, uint32_t* table) {
  const uint32_t q = threadIdx.x;  // 0..15: E[q]; 16..19: D[q - 16]
  qm31 b{{1u, 0u, 0u, 0u}};
  if (q < 16) {
    const uint32_t bits[4] = {0, 1, 10, 11};
    for (uint32_t k = 0; k < 4; k++)
      if ((q >> k) & 1u) b = qm_mul(b, ld_factor(factors, log_n, bits[k]));
  } else if (q < 20) {
    for (uint32_t k = 0; k < 2; k++)
      if (((q - 16) >> k) & 1u) b = qm_mul(b, ld_factor(factors, log_n, 12 + k));
  }
  if (q < 20) reinterpret_cast<qm31*>(table)[q] = b;
}

__global__ __launch_bounds__(EV_WG) void k_eval_big(const uint32_t* __restrict__ c, const uint32_t* table,
                                                    const uint32_t* factors, uint32_t log_n,
                                                    uint32_t* __restrict__ part) {
  __shared__ qm31 q[EV_WG];
  const uint32_t t = threadIdx.x;
  const uint4* p = reinterpret_cast<const uint4*>(c) + ((uint64_t)blockIdx.x << (EV_BIG_LOG - 2)) + t;
  uint4 x[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) x[m] = p[m * EV_WG];
  uniform_t<qm31>* E = (uniform_t<qm31>*)table;
  qm31 w[4];
#pragma unroll
  for (uint32_t m23 = 0; m23 < 4; m23++) {
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      uint32_t r = 0;
#pragma unroll
      for (uint32_t m01 = 0; m01 < 4; m01++) {
        const uint4 v = x[m01 + 4 * m23];
        // 4 products < 4 P^2 < 2^64
        const uint64_t acc = (uint64_t)v.x * E[4 * m01].v[k] + (uint64_t)v.y * E[4 * m01 + 1].v[k] +
                             (uint64_t)v.z * E[4 * m01 + 2].v[k] + (uint64_t)v.w * E[4 * m01 + 3].v[k];
        r = m31_add(r, m31_reduce64(acc));
      }
      w[m23].v[k] = r;
    }
  }
  qm31 v = w[0];
#pragma unroll
  for (uint32_t m23 = 1; m23 < 4; m23++)
    v = qm_add(v, qm_mul(ld_uniform(reinterpret_cast<const qm31*>(table) + 16 + m23), w[m23]));
  const qm31 s = lds_fold(q, v, factors, log_n, 2);
  if (t == 0) reinterpret_cast<qm31*>(part)[blockIdx.x] = s;
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
  uint32_t groups_log, fb;
  uint32_t* in = s1;
  uint32_t* out = s2;
  if (log_n >= EV_BIG_LOG) {
    // the 20-entry table sits in s2 until the first tree level overwrites it
    hipLaunchKernelGGL(k_eval_setup, dim3(1), dim3(64), 0, s, factors, log_n, s2);
    VX_HIP(hipGetLastError());
    groups_log = log_n - EV_BIG_LOG;
    hipLaunchKernelGGL(k_eval_big, dim3(1u << groups_log), dim3(EV_WG), 0, s, c, s2, factors, log_n, s1);
    VX_HIP(hipGetLastError());
    fb = EV_BIG_LOG;
  } else {
    const uint32_t lanes_log = log_n - EV_LANE_LOG;
    const uint32_t wg_log = lanes_log < EV_WG_LOG ? lanes_log : EV_WG_LOG;
    groups_log = lanes_log - wg_log;
    hipLaunchKernelGGL(k_eval_blocks, dim3(1u << groups_log), dim3(1u << wg_log), 0, s, c, factors, log_n, s1);
    VX_HIP(hipGetLastError());
    fb = EV_LANE_LOG + wg_log;
  }
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
// Many columns at one point (the OODS step evaluates every committed column
// at z; cuda/batch_eval_at_point.cu:11-100 restricts this to n <= 64). A
// column of 2^L coefficients uses the last L of the max_log_n factors, so
// coefficient bit b of every column is folded by the same
// F[b] = factors[max_log_n - 1 - b]. Columns are cut into chunks of 2^14
// coefficients (shorter columns are one chunk, the words past their end
// read as zero, which no factor can change), and one launch runs every
// chunk of every column with the k_eval_big arithmetic (work list: each
// workgroup binary-searches its column). Per column, the chunk sums are
// then folded 256 at a time with F[14..], one launch per level for all
// columns.
struct EvalChunkJob {
  const uint32_t* col;
  uint32_t log_n;
  uint32_t first_chunk;  // jobs sorted by first_chunk
};
struct EvalFoldJob {
  uint32_t in;   // first QM31 of the group in the input level
  uint32_t out;  // its QM31 in the output level
  uint32_t n;    // sums in the group (a power of two, <= 256)
  uint32_t fb;   // coefficient bit folded by the group's level 0
};

__device__ __forceinline__ qm31 factor_or_one(const uint32_t* factors, uint32_t max_log_n, uint32_t b) {
  return b < max_log_n ? ld_factor(factors, max_log_n, b) : qm31{{1u, 0u, 0u, 0u}};
}

// table: E[16] over bits {0,1,10,11}, D[4] over {12,13}, F[2..9] at 20..27
__global__ void k_eval_batch_setup(const uint32_t* factors, uint32_t max_log_n, uint32_t* table) {
  const uint32_t q = threadIdx.x;
  qm31 b{{1u, 0u, 0u, 0u}};
  if (q < 16) {
    const uint32_t bits[4] = {0, 1, 10, 11};
    for (uint32_t k = 0; k < 4; k++)
      if ((q >> k) & 1u) b = qm_mul(b, factor_or_one(factors, max_log_n, bits[k]));
  } else if (q < 20) {
    for (uint32_t k = 0; k < 2; k++)
      if (((q - 16) >> k) & 1u) b = qm_mul(b, factor_or_one(factors, max_log_n, 12 + k));
  } else if (q < 28) {
    b = factor_or_one(factors, max_log_n, 2 + (q - 20));
  } else {
    return;
  }
  reinterpret_cast<qm31*>(table)[q] = b;
}

// 256-lane LDS fold of q[0..256) with the factors f[0..8) (uniform).
__device__ qm31 lds_fold_table(qm31* q, qm31 v, uniform_t<qm31>* f) {
  const uint32_t t = threadIdx.x;
  q[t] = v;
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < EV_WG_LOG; j++) {
    const qm31 fj{{f[j].v[0], f[j].v[1], f[j].v[2], f[j].v[3]}};
    if ((t & ((2u << j) - 1)) == 0) q[t] = qm_add(q[t], qm_mul(fj, q[t + (1u << j)]));
    __syncthreads();
  }
  return q[0];
}

// NP points (1, or 2 for the OODS step's z and z_next) from one read of the
// coefficients: table holds NP 28-entry tables, part NP arrays of
// part_stride QM31.
template <uint32_t NP>
__global__ __launch_bounds__(EV_WG) void k_eval_batch_chunks(const EvalChunkJob* jobs, uint32_t n_jobs,
                                                             const uint32_t* table, uint32_t* part,
                                                             uint32_t part_stride) {
  __shared__ qm31 q[EV_WG];
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ld_uniform(&jobs[mid].first_chunk) <= b) lo = mid;
    else hi = mid;
  }
  const uint32_t* col = ld_uniform(&jobs[lo].col);
  const uint32_t L = ld_uniform(&jobs[lo].log_n);
  const uint64_t chunk = b - ld_uniform(&jobs[lo].first_chunk);
  const uint64_t n = 1ull << L;
  uint4 x[16];
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) {
    const uint64_t w = (chunk << EV_BIG_LOG) + 4ull * (t + m * EV_WG);  // first word of the load
    if (L >= 2) {
      x[m] = w < n ? reinterpret_cast<const uint4*>(col)[w >> 2] : make_uint4(0u, 0u, 0u, 0u);
    } else {
      x[m] = make_uint4(w < n ? col[w] : 0u, w + 1 < n ? col[w + 1] : 0u, 0u, 0u);
    }
  }
#pragma unroll
  for (uint32_t pt = 0; pt < NP; pt++) {
    const uint32_t* tab = table + 28 * 4 * pt;
    uniform_t<qm31>* E = (uniform_t<qm31>*)tab;
    qm31 w4[4];
#pragma unroll
    for (uint32_t m23 = 0; m23 < 4; m23++) {
#pragma unroll
      for (uint32_t k = 0; k < 4; k++) {
        uint32_t r = 0;
#pragma unroll
        for (uint32_t m01 = 0; m01 < 4; m01++) {
          const uint4 v = x[m01 + 4 * m23];
          const uint64_t acc = (uint64_t)v.x * E[4 * m01].v[k] + (uint64_t)v.y * E[4 * m01 + 1].v[k] +
                               (uint64_t)v.z * E[4 * m01 + 2].v[k] + (uint64_t)v.w * E[4 * m01 + 3].v[k];
          r = m31_add(r, m31_reduce64(acc));
        }
        w4[m23].v[k] = r;
      }
    }
    qm31 v = w4[0];
#pragma unroll
    for (uint32_t m23 = 1; m23 < 4; m23++)
      v = qm_add(v, qm_mul(ld_uniform(reinterpret_cast<const qm31*>(tab) + 16 + m23), w4[m23]));
    const qm31 sum = lds_fold_table(q, v, (uniform_t<qm31>*)tab + 20);
    if (t == 0) reinterpret_cast<qm31*>(part)[(uint64_t)pt * part_stride + b] = sum;
    if (NP > 1) __syncthreads();  // q is reused by the next point's fold
  }
}

// blockIdx.y = point: its factors, input and output arrays lie f_stride,
// in_stride and out_stride words after the previous point's.
__global__ __launch_bounds__(EV_WG) void k_eval_batch_fold(const EvalFoldJob* jobs, const uint32_t* factors,
                                                           uint32_t max_log_n, const uint32_t* in, uint32_t* out,
                                                           uint32_t f_stride, uint32_t in_stride,
                                                           uint32_t out_stride) {
  __shared__ qm31 q[EV_WG];
  const uint32_t t = threadIdx.x;
  factors += (uint64_t)blockIdx.y * f_stride;
  in += (uint64_t)blockIdx.y * in_stride;
  out += (uint64_t)blockIdx.y * out_stride;
  uniform_t<EvalFoldJob>& J = as_uniform(jobs + blockIdx.x);
  const uint32_t n = J.n, fb = J.fb;
  qm31 f[EV_WG_LOG];
#pragma unroll
  for (uint32_t j = 0; j < EV_WG_LOG; j++) f[j] = factor_or_one(factors, max_log_n, fb + j);
  q[t] = t < n ? reinterpret_cast<const qm31*>(in)[J.in + t] : qm31{{0u, 0u, 0u, 0u}};
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < EV_WG_LOG; j++) {
    if ((t & ((2u << j) - 1)) == 0) q[t] = qm_add(q[t], qm_mul(f[j], q[t + (1u << j)]));
    __syncthreads();
  }
  if (t == 0) reinterpret_cast<qm31*>(out)[J.out] = q[0];
}

// result[k] = QM31 at src[k] (device pointers, one per column)
__global__ void k_copy_qm31(const uint32_t* const* src, uint32_t* results, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  reinterpret_cast<qm31*>(results)[k] = *reinterpret_cast<const qm31*>(src[k]);
}

// np points (1 or 2): factors holds np tables of 4 max_log_n words back to
// back (4 words when max_log_n is 0), results np arrays of n_cols QM31.
static void eval_batch_points(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                              const uint32_t* factors, uint32_t np, uint32_t max_log_n, uint32_t* results,
                              hipStream_t s) {
  if (n_cols == 0) return;
  require(np == 1 || np == 2, "eval_at_point_batch: one or two points");
  const uint32_t f_stride = 4 * (max_log_n ? max_log_n : 1);
  require(max_log_n <= 31, "eval_at_point_batch: max_log_n must be <= 31");
  std::vector<EvalChunkJob> jobs(n_cols);
  std::vector<uint32_t> count(n_cols), off(n_cols);
  uint64_t total = 0;
  for (uint32_t k = 0; k < n_cols; k++) {
    const uint32_t L = h_log_sizes[k];
    require(h_cols[k] != nullptr && L <= max_log_n, "eval_at_point_batch: null column or log size above max_log_n");
    require(L < 2 || (reinterpret_cast<uintptr_t>(h_cols[k]) & 15u) == 0,
            "eval_at_point_batch: columns must be 16-byte aligned");
    count[k] = L > EV_BIG_LOG ? 1u << (L - EV_BIG_LOG) : 1u;
    off[k] = (uint32_t)total;
    jobs[k] = {h_cols[k], L, (uint32_t)total};
    total += count[k];
  }
  require(total < (1ull << 31), "eval_at_point_batch: too many chunks");
  // every level's sums stay allocated until the results are copied: a column
  // that is down to one sum keeps it where it was written
  std::vector<DeviceBuf> keep;
  keep.emplace_back(sizeof(EvalChunkJob) * n_cols);
  VX_HIP(hipMemcpyAsync(keep.back().p, jobs.data(), sizeof(EvalChunkJob) * n_cols, hipMemcpyHostToDevice, s));
  const EvalChunkJob* d_jobs = static_cast<const EvalChunkJob*>(keep.back().p);
  keep.emplace_back(np * 28 * 16);
  uint32_t* d_table = keep.back().u32();
  keep.emplace_back(np * total * 16);
  uint32_t* level = keep.back().u32();
  uint32_t level_stride = 4 * (uint32_t)total;  // words between the points' arrays
  for (uint32_t pt = 0; pt < np; pt++)
    hipLaunchKernelGGL(k_eval_batch_setup, dim3(1), dim3(64), 0, s, factors + pt * f_stride, max_log_n,
                       d_table + 28 * 4 * pt);
  if (np == 2)
    hipLaunchKernelGGL(k_eval_batch_chunks<2>, dim3((uint32_t)total), dim3(EV_WG), 0, s, d_jobs, n_cols, d_table,
                       level, (uint32_t)total);
  else
    hipLaunchKernelGGL(k_eval_batch_chunks<1>, dim3((uint32_t)total), dim3(EV_WG), 0, s, d_jobs, n_cols, d_table,
                       level, (uint32_t)total);
  VX_HIP(hipGetLastError());
  // final_src[pt * n_cols + k]: where column k's sum for point pt ends up
  std::vector<const uint32_t*> final_src((size_t)np * n_cols, nullptr);
  auto settle = [&](uint32_t k, const uint32_t* lv, uint32_t stride) {
    for (uint32_t pt = 0; pt < np; pt++) final_src[(size_t)pt * n_cols + k] = lv + (uint64_t)pt * stride + 4ull * off[k];
  };
  for (uint32_t k = 0; k < n_cols; k++)
    if (count[k] == 1) settle(k, level, level_stride);
  uint32_t fb = EV_BIG_LOG;
  for (;;) {
    std::vector<EvalFoldJob> fj;
    uint32_t next = 0;
    for (uint32_t k = 0; k < n_cols; k++) {
      if (count[k] <= 1) continue;
      const uint32_t w = count[k] < EV_WG ? count[k] : EV_WG;
      const uint32_t groups = count[k] / w;
      for (uint32_t g = 0; g < groups; g++) fj.push_back({off[k] + g * w, next + g, w, fb});
      off[k] = next;
      count[k] = groups;
      next += groups;
    }
    if (fj.empty()) break;
    keep.emplace_back(sizeof(EvalFoldJob) * fj.size());
    VX_HIP(hipMemcpyAsync(keep.back().p, fj.data(), sizeof(EvalFoldJob) * fj.size(), hipMemcpyHostToDevice, s));
    const EvalFoldJob* d_fj = static_cast<const EvalFoldJob*>(keep.back().p);
    keep.emplace_back((size_t)np * next * 16);
    uint32_t* out = keep.back().u32();
    hipLaunchKernelGGL(k_eval_batch_fold, dim3((uint32_t)fj.size(), np), dim3(EV_WG), 0, s, d_fj, factors,
                       max_log_n, level, out, f_stride, level_stride, 4 * next);
    VX_HIP(hipGetLastError());
    level = out;
    level_stride = 4 * next;
    for (uint32_t k = 0; k < n_cols; k++)
      if (count[k] == 1 && final_src[k] == nullptr) settle(k, level, level_stride);
    fb += EV_WG_LOG;
  }
  const uint32_t n_out = np * n_cols;
  keep.emplace_back(sizeof(uint32_t*) * n_out);
  VX_HIP(hipMemcpyAsync(keep.back().p, final_src.data(), sizeof(uint32_t*) * n_out, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_copy_qm31, dim3((n_out + 255) / 256), dim3(256), 0, s,
                     static_cast<const uint32_t* const*>(keep.back().p), results, n_out);
  VX_HIP(hipGetLastError());
  // the staging buffers are freed on return: the caller's stream must be done
  VX_HIP(hipStreamSynchronize(s));
}

void eval_at_point_batch(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                         const uint32_t* factors, uint32_t max_log_n, uint32_t* results, hipStream_t s) {
  eval_batch_points(h_cols, h_log_sizes, n_cols, factors, 1, max_log_n, results, s);
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

// ------------------------------------------------------------- OODS step
// src/oods.rs:44-81: t drawn from the channel, z = ((1 - t^2) / (1 + t^2),
// 2t / (1 + t^2)) on the QM31 circle, z_next = z * step with step the
// generator of the 2^trace_log subgroup (circle.rs generator doubled
// 31 - trace_log times); then both points' folding factors for max_log:
// [z.y, z.x, 2 z.x^2 - 1, ...] (max_log entries) reversed.
__device__ void oods_factors(qm31 zx, qm31 zy, uint32_t max_log, uint32_t* out) {
  qm31 x = zx;
  for (uint32_t i = 0; i < max_log; i++) {
    const qm31 m = i == 0 ? zy : x;
    if (i > 0) {
      const qm31 x2 = qm_mul(x, x);
      x = qm_sub(qm_add(x2, x2), qm31{{1u, 0u, 0u, 0u}});
    }
    for (int k = 0; k < 4; k++) out[4 * (max_log - 1 - i) + k] = m.v[k];
  }
}

__global__ void k_oods_point(uint32_t* ch_words, uint32_t trace_log, uint32_t max_log, uint32_t* points, uint32_t* fz,
                             uint32_t* fzn) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Channel c;
  for (int i = 0; i < 8; i++) c.state[i] = ch_words[i];
  c.n_draws = ch_words[8];
  const qm31 t = ch_draw_felt(c);
  for (int i = 0; i < 8; i++) ch_words[i] = c.state[i];
  ch_words[8] = c.n_draws;
  const qm31 one{{1u, 0u, 0u, 0u}};
  const qm31 t2 = qm_mul(t, t);
  const qm31 inv = qm_inv(qm_add(one, t2));
  const qm31 zx = qm_mul(qm_sub(one, t2), inv), zy = qm_mul(qm_add(t, t), inv);
  const cpt st = c_rdouble(c_generator(), 31 - trace_log);
  const qm31 nx = qm_sub(qm_scale(zx, st.x), qm_scale(zy, st.y));
  const qm31 ny = qm_add(qm_scale(zx, st.y), qm_scale(zy, st.x));
  const qm31 pts[4] = {zx, zy, nx, ny};
  for (int p = 0; p < 4; p++)
    for (int k = 0; k < 4; k++) points[4 * p + k] = pts[p].v[k];
  oods_factors(zx, zy, max_log, fz);
  oods_factors(nx, ny, max_log, fzn);
}

// Channel::mix_felts (src/channel.rs:46-59) of the sampled values: the
// message is state || values, values being msg[8..] (msg[0..8) unused); it
// is a dependent chain of compressions, which the four lanes of a quad
// share (b2_compress_x4, about 440 instructions per lane and block instead
// of 960). One wavefront; lanes 0..3 write the new state.
__global__ __launch_bounds__(64) void k_oods_mix(uint32_t* ch_words, const uint32_t* msg, uint32_t n_felts) {
  if (blockIdx.x != 0 || n_felts == 0) return;
  const uint32_t q = threadIdx.x & 3u;
  const bool q0 = q & 1, q1 = q & 2;
  uint32_t h_lo = sel4(q0, q1, VX_B2_IV0 ^ 0x01010020u, VX_B2_IV1, VX_B2_IV2, VX_B2_IV3);
  uint32_t h_hi = sel4(q0, q1, VX_B2_IV4, VX_B2_IV5, VX_B2_IV6, VX_B2_IV7);
  const uint32_t nbytes = 32u + 16u * n_felts;
  for (uint32_t off = 0;; off += 64) {
    const uint32_t rem = nbytes - off;
    const bool last = rem <= 64;
    uint32_t m[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
      const uint32_t w = off / 4 + k;
      m[k] = 4 * k < rem ? (w < 8 ? ch_words[w] : msg[w]) : 0u;
    }
    b2_compress_x4(h_lo, h_hi, m, last ? nbytes : off + 64, last, q);
    if (last) break;
  }
  if (threadIdx.x < 4) {
    ch_words[q] = h_lo;
    ch_words[4 + q] = h_hi;
  }
  if (threadIdx.x == 0) ch_words[8] = 0;
}

void oods_sample(uint32_t* d_channel, const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                 uint32_t trace_log, uint32_t* d_points, uint32_t* d_values, hipStream_t s) {
  uint32_t max_log = 0;
  for (uint32_t k = 0; k < n_cols; k++) max_log = h_log_sizes[k] > max_log ? h_log_sizes[k] : max_log;
  const uint32_t fw = 4 * (max_log ? max_log : 1);
  DeviceBuf fac(8ull * fw), msg(4ull * (8 + 8ull * n_cols));
  uint32_t* fz = static_cast<uint32_t*>(fac.p);
  uint32_t* m = static_cast<uint32_t*>(msg.p);
  hipLaunchKernelGGL(k_oods_point, dim3(1), dim3(64), 0, s, d_channel, trace_log, max_log, d_points, fz, fz + fw);
  VX_HIP(hipGetLastError());
  if (n_cols) {
    // both points from one read of every column: values at z, then at z_next
    eval_batch_points(h_cols, h_log_sizes, n_cols, fz, 2, max_log, m + 8, s);
    hipLaunchKernelGGL(k_oods_mix, dim3(1), dim3(64), 0, s, d_channel, m, 2 * n_cols);
    VX_HIP(hipGetLastError());
    VX_HIP(hipMemcpyAsync(d_values, m + 8, 32ull * n_cols, hipMemcpyDeviceToDevice, s));
  }
  VX_HIP(hipStreamSynchronize(s));
}

}  // namespace vx
