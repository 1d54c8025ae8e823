// Kernels of the synthetic degree-2 multi-column AIR (definition in vx_d2.h):
// trace generation and the powers of the composition challenge. The
// quotient itself is produced inside the Merkle subtree kernel
// (LEAF_D2_QUOTIENT in merkle.hip), which also stores it for the circle fold.
#include "vx_d2.h"
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t D2_ROWS_PER_LANE = 4;

// Column blockIdx.y, rows [4 lane, 4 lane + 4); the 64-bit trace seed is
// read from device memory so a captured proof graph replays with new seeds.
__global__ void k_d2_trace(const uint32_t* seed_words, uint32_t* out, uint64_t n, uint64_t stride) {
  const uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * D2_ROWS_PER_LANE;
  if (first >= n) return;
  const uint64_t seed = (uint64_t)seed_words[0] | ((uint64_t)seed_words[1] << 32);
  const uint32_t c = blockIdx.y;
  uint32_t v[D2_ROWS_PER_LANE];
#pragma unroll
  for (uint32_t j = 0; j < D2_ROWS_PER_LANE; j++) v[j] = d2_value(seed, c, first + j);
  uint32_t* col = out + c * stride;
  if (first + D2_ROWS_PER_LANE <= n) {
    *reinterpret_cast<uint4*>(col + first) = make_uint4(v[0], v[1], v[2], v[3]);
  } else {
    for (uint32_t j = 0; first + j < n; j++) col[first + j] = v[j];
  }
}

void d2_trace(const uint32_t* seed_words, uint32_t* out, uint32_t n_cols, uint32_t log_n, uint64_t stride,
              hipStream_t s) {
  require(log_n >= 1 && log_n <= 31 && n_cols >= 1 && n_cols <= D2_MAX_COLS && stride >= (1ull << log_n),
          "d2_trace: bad shape");
  const uint64_t n = 1ull << log_n;
  const uint64_t lanes = (n + D2_ROWS_PER_LANE - 1) / D2_ROWS_PER_LANE;
  hipLaunchKernelGGL(k_d2_trace, dim3((uint32_t)((lanes + 255) / 256), n_cols), dim3(256), 0, s, seed_words, out,
                     n, stride);
  VX_HIP(hipGetLastError());
}

// One column (index `col` of the generator) of 2^log_n values.
__global__ void k_d2_trace_col(const uint32_t* seed_words, uint32_t* out, uint32_t col, uint64_t n) {
  const uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * D2_ROWS_PER_LANE;
  if (first >= n) return;
  const uint64_t seed = (uint64_t)seed_words[0] | ((uint64_t)seed_words[1] << 32);
#pragma unroll
  for (uint32_t j = 0; j < D2_ROWS_PER_LANE; j++)
    if (first + j < n) out[first + j] = d2_value(seed, col, first + j);
}

void d2_trace_col(const uint32_t* seed_words, uint32_t* out, uint32_t col, uint32_t log_n, hipStream_t s) {
  require(log_n >= 1 && log_n <= 31, "d2_trace_col: bad log_n");
  const uint64_t n = 1ull << log_n;
  const uint64_t lanes = (n + D2_ROWS_PER_LANE - 1) / D2_ROWS_PER_LANE;
  hipLaunchKernelGGL(k_d2_trace_col, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, s, seed_words, out, col, n);
  VX_HIP(hipGetLastError());
}

// Columns stored back to back: column c occupies [off[c], off[c+1]) (every
// length a multiple of 4) and holds the generator's column c. One launch
// for all columns; a lane finds its column by binary search over off[].
__global__ void k_d2_fill_columns(const uint32_t* seed_words, uint32_t* out, const uint64_t* off, uint32_t n_cols) {
  const uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * D2_ROWS_PER_LANE;
  if (first >= off[n_cols]) return;
  uint32_t lo = 0, hi = n_cols;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= first) lo = mid;
    else hi = mid;
  }
  const uint64_t seed = (uint64_t)seed_words[0] | ((uint64_t)seed_words[1] << 32);
  const uint64_t row = first - off[lo];
  uint32_t v[D2_ROWS_PER_LANE];
#pragma unroll
  for (uint32_t j = 0; j < D2_ROWS_PER_LANE; j++) v[j] = d2_value(seed, lo, row + j);
  *reinterpret_cast<uint4*>(out + first) = make_uint4(v[0], v[1], v[2], v[3]);
}

void d2_fill_columns(const uint32_t* seed_words, uint32_t* out, const uint64_t* d_off, uint32_t n_cols,
                     uint64_t total, hipStream_t s) {
  require(total % D2_ROWS_PER_LANE == 0 && n_cols >= 1, "d2_fill_columns: bad shape");
  const uint64_t lanes = total / D2_ROWS_PER_LANE;
  hipLaunchKernelGGL(k_d2_fill_columns, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, s, seed_words, out,
                     d_off, n_cols);
  VX_HIP(hipGetLastError());
}

// beta[k] = alpha^(k+1), k < K (one lane; K <= 1024 QM31 products).
__global__ void k_d2_powers(const uint32_t* alpha, uint32_t* beta, uint32_t k_cons) {
  if (threadIdx.x != 0) return;
  const qm31 a = {{alpha[0], alpha[1], alpha[2], alpha[3]}};
  qm31 p = a;
  for (uint32_t k = 0; k < k_cons; k++) {
#pragma unroll
    for (int c = 0; c < 4; c++) beta[4 * k + c] = p.v[c];
    p = qm_mul(p, a);
  }
}

void d2_powers(const uint32_t* alpha, uint32_t* beta, uint32_t k_cons, hipStream_t s) {
  require(k_cons >= 1 && k_cons <= 4096, "d2_powers: bad count");
  hipLaunchKernelGGL(k_d2_powers, dim3(1), dim3(64), 0, s, alpha, beta, k_cons);
  VX_HIP(hipGetLastError());
}

// Quotient weights of the K constraints: w[8k..8k+3] = beta_k c_k and
// w[8k+4..8k+7] = beta_k d_k with beta_k = alpha^(k+1) and (c_k, d_k) from the
// constraint table (8 words per constraint, c and d first). One lane.
__global__ void k_d2_weights(const uint32_t* alpha, const uint32_t* cons, uint32_t* w, uint32_t k_cons) {
  if (threadIdx.x != 0) return;
  const qm31 a = {{alpha[0], alpha[1], alpha[2], alpha[3]}};
  qm31 p = a;
  for (uint32_t k = 0; k < k_cons; k++) {
    const qm31 g = qm_scale(p, cons[8 * k]), d = qm_scale(p, cons[8 * k + 1]);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      w[8 * k + c] = g.v[c];
      w[8 * k + 4 + c] = d.v[c];
    }
    p = qm_mul(p, a);
  }
}

void d2_weights(const uint32_t* alpha, const uint32_t* cons, uint32_t* w, uint32_t k_cons, hipStream_t s) {
  require(k_cons >= 1 && k_cons <= D2_MAX_CONS, "d2_weights: bad count");
  hipLaunchKernelGGL(k_d2_weights, dim3(1), dim3(64), 0, s, alpha, cons, w, k_cons);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
