// Cairo VM AIR quotient kernels (cuda/cairo_constraint.cu): the quotient in
// its full, slab and chunked forms (one row per lane, vx_cairo.h), the
// vanishing polynomial of the trace domain at every evaluation position, and
// y += s x. The fused quotient + Merkle commitment is the LEAF_CAIRO_QUOTIENT
// producer of merkle.hip.
#include "vx_cairo.h"

namespace vx {

static constexpr uint32_t CA_THREADS = 256;

// Every column is read at consecutive rows by consecutive lanes (one
// 256-byte run per wavefront and column); the next-row reads hit the same
// lines blowup_step rows later. Alphas and challenges are wave-uniform
// (scalar loads).
__global__ __launch_bounds__(CA_THREADS) void k_cairo_quotient(CairoArgs a, uint64_t rows) {
  const uint64_t r = (uint64_t)blockIdx.x * CA_THREADS + threadIdx.x;
  if (r >= rows) return;
  const qm31 q = cairo_row(a, r);
  const uint64_t o = a.o_base + r;
#pragma unroll
  for (int k = 0; k < 4; k++) a.out[k][o] = q.v[k];
}

void cairo_quotient(const CairoArgs& a, uint64_t rows, hipStream_t s) {
  if (!rows) return;
  const uint64_t grid = (rows + CA_THREADS - 1) / CA_THREADS;
  require(grid < (1ull << 31), "cairo_quotient: too many rows");
  hipLaunchKernelGGL(k_cairo_quotient, dim3((uint32_t)grid), dim3(CA_THREADS), 0, s, a, rows);
  VX_HIP(hipGetLastError());
}

// The point init * step^bitrev(i): step^j as the product of the precomputed
// step^(2^b) over the set bits b of j (kernel argument table).
struct StepPow {
  uint32_t x[31], y[31];
};

__global__ __launch_bounds__(CA_THREADS) void k_cairo_vanishing(uint32_t ix, uint32_t iy, StepPow sp, uint32_t* out,
                                                                uint32_t log_eval, uint32_t log_n, int inverse) {
  const uint64_t i = (uint64_t)blockIdx.x * CA_THREADS + threadIdx.x;
  if (i >= (1ull << log_eval)) return;
  const uint32_t j = log_eval ? __builtin_bitreverse32((uint32_t)i) >> (32 - log_eval) : 0;
  uint32_t px = ix, py = iy;
  for (uint32_t b = 0; b < log_eval; b++) {
    if ((j >> b) & 1u) {
      const uint32_t nx = m31_sub(m31_mul(px, sp.x[b]), m31_mul(py, sp.y[b]));
      py = m31_add(m31_mul(px, sp.y[b]), m31_mul(py, sp.x[b]));
      px = nx;
    }
  }
  // Z_H(x) = f_{log_n}(x) + 1, f_0 = x, f_{k+1} = 2 f_k^2 - 1
  uint32_t v = px;
  for (uint32_t k = 0; k < log_n; k++) {
    const uint32_t sq = m31_mul(v, v);
    v = m31_sub(m31_add(sq, sq), 1);
  }
  const uint32_t z = m31_add(v, 1);
  out[i] = inverse ? m31_inv(z) : z;
}

void cairo_vanishing(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* out, uint32_t log_eval,
                     uint32_t log_n, bool inverse, hipStream_t s) {
  require(out != nullptr && log_eval <= 31 && log_n <= 31, "vanishing: bad arguments");
  StepPow sp{};
  uint32_t x = sx, y = sy;
  for (uint32_t b = 0; b < 31; b++) {
    sp.x[b] = x;
    sp.y[b] = y;
    const uint32_t nx = m31_sub(m31_mul(x, x), m31_mul(y, y));
    y = m31_add(m31_mul(x, y), m31_mul(y, x));
    x = nx;
  }
  const uint64_t n = 1ull << log_eval;
  hipLaunchKernelGGL(k_cairo_vanishing, dim3((uint32_t)((n + CA_THREADS - 1) / CA_THREADS)), dim3(CA_THREADS), 0, s,
                     ix, iy, sp, out, log_eval, log_n, inverse ? 1 : 0);
  VX_HIP(hipGetLastError());
}

__global__ __launch_bounds__(CA_THREADS) void k_axpy_m31(uint32_t sc, const uint32_t* x, uint32_t* y, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * CA_THREADS + threadIdx.x;
  if (i < n) y[i] = m31_add(y[i], m31_mul(sc, x[i]));
}

void axpy_m31(uint32_t sc, const uint32_t* x, uint32_t* y, uint64_t n, hipStream_t s) {
  if (!n) return;
  require(x && y, "axpy: null pointer");
  hipLaunchKernelGGL(k_axpy_m31, dim3((uint32_t)((n + CA_THREADS - 1) / CA_THREADS)), dim3(CA_THREADS), 0, s, sc, x,
                     y, n);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
