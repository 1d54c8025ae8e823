// Trace generation, Fibonacci quotient and FRI fold kernels.
//
// Reference behaviour: src/air.rs:30-45 (trace), cuda/constraint.cu:10-64
// (quotient), cuda/fri.cu:58-126 (folds). These are the stand-alone entry
// points of the C-ABI; the prover fuses quotient and folds into the Merkle
// subtree kernel instead (merkle.hip).
#include "vx_internal.h"

namespace vx {

// t_k = a F(k-1) + b F(k) with the standard Fibonacci F; each lane jumps to
// its first index with fast doubling (F(2m) = F(m)(2F(m+1) - F(m)),
// F(2m+1) = F(m)^2 + F(m+1)^2), then steps t_{k+2} = t_{k+1} + t_k. Exact
// arithmetic, so this equals the sequential recurrence.
static constexpr uint32_t TRACE_PER_LANE = 16;

__device__ __forceinline__ void fib_pair(uint64_t k, uint32_t& fk, uint32_t& fk1) {
  uint32_t f = 0, g = 1;  // F(m), F(m+1) with m = 0
  for (int bit = 31; bit >= 0; bit--) {  // k < 2^32
    uint32_t c = m31_mul(f, m31_sub(m31_add(g, g), f));
    uint32_t d = m31_add(m31_mul(f, f), m31_mul(g, g));
    if ((k >> bit) & 1) {
      f = d;
      g = m31_add(c, d);
    } else {
      f = c;
      g = d;
    }
  }
  fk = f;
  fk1 = g;
}

__device__ __forceinline__ void fib_fill(uint32_t a, uint32_t b, uint32_t* out, uint64_t n) {
  const uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * TRACE_PER_LANE;
  if (first >= n) return;
  uint32_t fk, fk1;
  fib_pair(first, fk, fk1);
  uint32_t fkm1 = m31_sub(fk1, fk);  // F(k-1)
  uint32_t t0 = m31_add(m31_mul(a, fkm1), m31_mul(b, fk));
  uint32_t t1 = m31_add(m31_mul(a, fk), m31_mul(b, fk1));
  uint32_t v[TRACE_PER_LANE];
  v[0] = t0;
  v[1] = t1;
#pragma unroll
  for (uint32_t i = 2; i < TRACE_PER_LANE; i++) v[i] = m31_add(v[i - 1], v[i - 2]);
  if (first + TRACE_PER_LANE <= n) {
#pragma unroll
    for (uint32_t i = 0; i < TRACE_PER_LANE; i += 4)
      *reinterpret_cast<uint4*>(out + first + i) = make_uint4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  } else {
    for (uint32_t i = 0; first + i < n; i++) out[first + i] = v[i];
  }
}

__global__ void k_fib_trace(uint32_t a, uint32_t b, uint32_t* out, uint64_t n) { fib_fill(a, b, out, n); }

// Public inputs read from device memory, so a captured graph replays with new (a, b).
__global__ void k_fib_trace_dev(const uint32_t* ab, uint32_t* out, uint64_t n) { fib_fill(ab[0], ab[1], out, n); }

static uint32_t trace_blocks(uint64_t n) {
  uint64_t lanes = (n + TRACE_PER_LANE - 1) / TRACE_PER_LANE;
  return (uint32_t)((lanes + 255) / 256);
}

void fib_trace(uint32_t a, uint32_t b, uint32_t* out, uint32_t log_n, hipStream_t s) {
  require(log_n >= 1 && log_n <= 31 && a < P && b < P, "fib_trace: bad arguments");
  uint64_t n = 1ull << log_n;
  hipLaunchKernelGGL(k_fib_trace, dim3(trace_blocks(n)), dim3(256), 0, s, a, b, out, n);
  VX_HIP(hipGetLastError());
}

void fib_trace_dev(const uint32_t* ab, uint32_t* out, uint32_t log_n, hipStream_t s) {
  require(log_n >= 1 && log_n <= 31, "fib_trace: bad log_n");
  uint64_t n = 1ull << log_n;
  hipLaunchKernelGGL(k_fib_trace_dev, dim3(trace_blocks(n)), dim3(256), 0, s, ab, out, n);
  VX_HIP(hipGetLastError());
}

// inputs[0..1] = (a, b) and the proof header words that carry them
// (Fibonacci: the public inputs; multi-column AIR: the trace seed words).
__global__ void k_set_inputs(uint32_t* inputs, uint32_t* proof, uint32_t word_a, uint32_t a, uint32_t b) {
  if (threadIdx.x == 0) {
    inputs[0] = a;
    inputs[1] = b;
    proof[word_a] = a;
    proof[word_a + 1] = b;
  }
}

void set_inputs(uint32_t* inputs, uint32_t* proof, uint32_t word_a, uint32_t a, uint32_t b, hipStream_t s) {
  hipLaunchKernelGGL(k_set_inputs, dim3(1), dim3(64), 0, s, inputs, proof, word_a, a, b);
  VX_HIP(hipGetLastError());
}

// q_i = alpha (t[(i+2) % n] - (t[(i+1) % n] + t[i])) for i in [offset, offset + chunk).
__global__ void k_fib_quotient(const uint32_t* t, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3, qm31 alpha,
                               uint64_t offset, uint64_t chunk, uint64_t n) {
  uint64_t li = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= chunk) return;
  uint64_t i = offset + li;
  uint64_t i1 = i + 1 >= n ? i + 1 - n : i + 1;
  uint64_t i2 = i + 2 >= n ? i + 2 - n : i + 2;
  uint32_t c = m31_sub(t[i2], m31_add(t[i1], t[i]));
  qm31 r = qm_scale(alpha, c);
  o0[li] = r.v[0];
  o1[li] = r.v[1];
  o2[li] = r.v[2];
  o3[li] = r.v[3];
}
void fib_quotient(const uint32_t* t, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3, qm31 alpha,
                  uint64_t offset, uint64_t chunk, uint64_t n, hipStream_t s) {
  require(n >= 2 && offset + chunk <= n, "fib_quotient: chunk outside the domain");
  if (chunk == 0) return;
  hipLaunchKernelGGL(k_fib_quotient, dim3((uint32_t)((chunk + 255) / 256)), dim3(256), 0, s, t, o0, o1, o2, o3,
                     alpha, offset, chunk, n);
  VX_HIP(hipGetLastError());
}

struct Cols4 {
  const uint32_t* c[4];
};
struct MCols4 {
  uint32_t* c[4];
};

__global__ void k_fold_line(Cols4 in, const uint32_t* tw, MCols4 out, qm31 alpha, uint64_t half_n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  qm31 f0 = {{in.c[0][2 * i], in.c[1][2 * i], in.c[2][2 * i], in.c[3][2 * i]}};
  qm31 f1 = {{in.c[0][2 * i + 1], in.c[1][2 * i + 1], in.c[2][2 * i + 1], in.c[3][2 * i + 1]}};
  qm31 r = fri_fold_pair(f0, f1, alpha, tw[i]);
  for (int k = 0; k < 4; k++) out.c[k][i] = r.v[k];
}
void fold_line(const uint32_t* const in[4], const uint32_t* tw, uint32_t* const out[4], qm31 alpha, uint64_t half_n,
               hipStream_t s) {
  if (half_n == 0) return;
  Cols4 a{{in[0], in[1], in[2], in[3]}};
  MCols4 o{{out[0], out[1], out[2], out[3]}};
  hipLaunchKernelGGL(k_fold_line, dim3((uint32_t)((half_n + 255) / 256)), dim3(256), 0, s, a, tw, o, alpha, half_n);
  VX_HIP(hipGetLastError());
}

// dst_i = dst_i * alpha^2 + fold(src_2i, src_2i+1)  (cuda/fri.cu:92-126)
__global__ void k_fold_circle(MCols4 dst, Cols4 src, const uint32_t* tw, qm31 alpha, qm31 alpha_sq, uint64_t half_n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  qm31 f0 = {{src.c[0][2 * i], src.c[1][2 * i], src.c[2][2 * i], src.c[3][2 * i]}};
  qm31 f1 = {{src.c[0][2 * i + 1], src.c[1][2 * i + 1], src.c[2][2 * i + 1], src.c[3][2 * i + 1]}};
  qm31 prev = {{dst.c[0][i], dst.c[1][i], dst.c[2][i], dst.c[3][i]}};
  qm31 r = qm_add(qm_mul(prev, alpha_sq), fri_fold_pair(f0, f1, alpha, tw[i]));
  for (int k = 0; k < 4; k++) dst.c[k][i] = r.v[k];
}
void fold_circle_into_line(uint32_t* const dst[4], const uint32_t* const src[4], const uint32_t* tw, qm31 alpha,
                           qm31 alpha_sq, uint64_t half_n, hipStream_t s) {
  if (half_n == 0) return;
  MCols4 d{{dst[0], dst[1], dst[2], dst[3]}};
  Cols4 a{{src[0], src[1], src[2], src[3]}};
  hipLaunchKernelGGL(k_fold_circle, dim3((uint32_t)((half_n + 255) / 256)), dim3(256), 0, s, d, a, tw, alpha,
                     alpha_sq, half_n);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
