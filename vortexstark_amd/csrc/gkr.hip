// GKR sum-check operations of the stwo plugin (cuda/gkr.cu, driven by
// crates/cuda-backend/src/gkr_ops.rs): folding an MLE's first variable,
// the eq table, the next layer of a grand-product or LogUp circuit, and the
// two sums of the round polynomial. Secure columns hold QM31 element i in
// words 4i..4i+3 (the reference's CudaColumn<SecureField>), base columns one
// M31 word per element.
//
// The element-wise kernels move 16-byte elements with one vector access
// each when the columns are 16-byte aligned (word accesses otherwise) and
// are bound by HBM; products are 4x4 lazy dot products (qm_mul_lazy). The
// sums accumulate canonical values in 64 bits per lane, reduce each lane
// once, then across the wavefront and the workgroup; a second launch folds
// the workgroups' partial sums into the first block, so the caller reads
// one block.
#include <atomic>

#include "vx_internal.h"

namespace vx {

namespace {

constexpr uint32_t GKR_THREADS = 256;

template <bool A16>
__device__ __forceinline__ qm31 ldqe(const uint32_t* p, uint64_t i) {
  if constexpr (A16) {
    const uint4 x = ld4_global(p + 4 * i);
    return {{x.x, x.y, x.z, x.w}};
  } else {
    const uint32_t* q = p + 4 * i;
    return {{q[0], q[1], q[2], q[3]}};
  }
}
template <bool A16>
__device__ __forceinline__ void stqe(uint32_t* p, uint64_t i, const qm31& x) {
  if constexpr (A16) {
    *reinterpret_cast<uint4*>(p + 4 * i) = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  } else {
    uint32_t* q = p + 4 * i;
    q[0] = x.v[0];
    q[1] = x.v[1];
    q[2] = x.v[2];
    q[3] = x.v[3];
  }
}
// A QM31 argument (4 words, any alignment) read into scalar registers.
__device__ __forceinline__ qm31 ldq_arg(const uint32_t* p) {
  return {{ld_uniform(p), ld_uniform(p + 1), ld_uniform(p + 2), ld_uniform(p + 3)}};
}
__device__ __forceinline__ qm31 qm_one_minus(const qm31& y) {
  return {{m31_sub(1u, y.v[0]), m31_neg(y.v[1]), m31_neg(y.v[2]), m31_neg(y.v[3])}};
}

// out[i] = r (in[i + h] - in[i]) + in[i]  (cuda/gkr.cu:20-45)
template <bool A16>
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_fix_base(const uint32_t* in, uint32_t* out, const uint32_t* r,
                                                             uint64_t h) {
  const uint64_t i = (uint64_t)blockIdx.x * GKR_THREADS + threadIdx.x;
  if (i >= h) return;
  const qm31 rq = ldq_arg(r);
  const uint32_t a = in[i];
  qm31 o = qm_scale(rq, m31_sub(in[i + h], a));
  o.v[0] = m31_add(o.v[0], a);
  stqe<A16>(out, i, o);
}

// The same over a secure column (cuda/gkr.cu:47-74). Element i is read and
// written only by lane i, so out may be in.
template <bool A16>
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_fix_secure(const uint32_t* in, uint32_t* out, const uint32_t* r,
                                                               uint64_t h) {
  const uint64_t i = (uint64_t)blockIdx.x * GKR_THREADS + threadIdx.x;
  if (i >= h) return;
  const qm31 rq = ldq_arg(r);
  const qm31 a = ldqe<A16>(in, i), b = ldqe<A16>(in, i + h);
  stqe<A16>(out, i, qm_add(qm_mul_lazy(rq, qm_sub(b, a)), a));
}

__global__ void k_gkr_eq_init(uint32_t* buf, const uint32_t* v) {
  if (threadIdx.x < 4) buf[threadIdx.x] = v[threadIdx.x];
}

// One doubling pass (cuda/gkr.cu:84-118): t = buf[j] y, buf[cur + j] = t,
// buf[j] -= t, j < cur.
template <bool A16>
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_eq_pass(uint32_t* buf, const uint32_t* y, uint64_t cur) {
  const uint64_t j = (uint64_t)blockIdx.x * GKR_THREADS + threadIdx.x;
  if (j >= cur) return;
  const qm31 v = ldqe<A16>(buf, j), t = qm_mul_lazy(v, ldq_arg(y));
  stqe<A16>(buf, cur + j, t);
  stqe<A16>(buf, j, qm_sub(v, t));
}

// The whole eq table in one launch. Entry x of the reference's passes is
// v prod_i (bit i of x ? y[k-1-i] : 1 - y[k-1-i]), a product that splits
// over the index bits: each workgroup owns GKR_EQ_HPW consecutive values H
// of the high bits (2^kl entries each), builds the low-bit factor table in
// LDS by the same doubling passes (from 1), forms v times the high-bit
// factors of its H values (one lane each), and writes their products, so
// every entry costs one product and one 16-byte store. Field products are
// exact, so the entries equal the pass-by-pass ones.
constexpr uint32_t GKR_EQ_LO_MAX = 10;
constexpr uint32_t GKR_EQ_HPW = 16;
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_eq_table(uint32_t* buf, const uint32_t* y, uint32_t k,
                                                             uint32_t kl, uint32_t hpw, const uint32_t* v, bool a16) {
  __shared__ uint4 lo[1u << GKR_EQ_LO_MAX];
  __shared__ uint4 hi[GKR_EQ_HPW];
  const uint32_t t = threadIdx.x;
  const uint64_t H0 = (uint64_t)blockIdx.x * hpw;
  if (t < hpw) {
    const uint64_t H = H0 + t;
    qm31 acc = ldq_arg(v);
    for (uint32_t i = kl; i < k; i++) {
      const qm31 yi = ldq_arg(y + 4 * (k - 1 - i));
      acc = qm_mul_lazy(acc, ((H >> (i - kl)) & 1) ? yi : qm_one_minus(yi));
    }
    hi[t] = make_uint4(acc.v[0], acc.v[1], acc.v[2], acc.v[3]);
  }
  if (t == 0) lo[0] = make_uint4(1u, 0u, 0u, 0u);
  __syncthreads();
  for (uint32_t i = 0; i < kl; i++) {
    const uint32_t cur = 1u << i;
    const qm31 yi = ldq_arg(y + 4 * (k - 1 - i));
    for (uint32_t j = t; j < cur; j += GKR_THREADS) {
      const uint4 e = lo[j];
      const qm31 ev{{e.x, e.y, e.z, e.w}}, tt = qm_mul_lazy(ev, yi), d = qm_sub(ev, tt);
      lo[cur + j] = make_uint4(tt.v[0], tt.v[1], tt.v[2], tt.v[3]);
      lo[j] = make_uint4(d.v[0], d.v[1], d.v[2], d.v[3]);
    }
    __syncthreads();
  }
  const uint32_t per = hpw << kl;
  for (uint32_t idx = t; idx < per; idx += GKR_THREADS) {
    const uint4 h = hi[idx >> kl], e = lo[idx & ((1u << kl) - 1)];
    const qm31 o = qm_mul_lazy({{h.x, h.y, h.z, h.w}}, {{e.x, e.y, e.z, e.w}});
    if (a16) stqe<true>(buf, (H0 << kl) + idx, o);
    else stqe<false>(buf, (H0 << kl) + idx, o);
  }
}

enum : int { GKR_PROD = 0, GKR_GENERIC = 1, GKR_MULT = 2, GKR_SINGLES = 3 };

// Numerator of element e of a layer: QM31 (generic), M31 embedded (mult),
// or 1 (singles).
template <int KIND, bool A16>
__device__ __forceinline__ qm31 ld_num(const uint32_t* num, uint64_t e) {
  if constexpr (KIND == GKR_MULT) return {{num[e], 0u, 0u, 0u}};
  else if constexpr (KIND == GKR_SINGLES) return {{1u, 0u, 0u, 0u}};
  else return ldqe<A16>(num, e);
}

// Fraction sum of a pair (n0/d0 + n1/d1 = (n0 d1 + n1 d0) / (d0 d1)); the
// M31 and unit numerators take the cheaper forms of cuda/gkr.cu:160-212.
template <int KIND>
__device__ __forceinline__ qm31 frac_num(const qm31& n0, const qm31& d0, const qm31& n1, const qm31& d1) {
  if constexpr (KIND == GKR_MULT) return qm_add(qm_scale(d1, n0.v[0]), qm_scale(d0, n1.v[0]));
  else if constexpr (KIND == GKR_SINGLES) return qm_add(d0, d1);
  else return qm_add(qm_mul_lazy(n0, d1), qm_mul_lazy(n1, d0));
}

// next_layer (cuda/gkr.cu:120-212): element i of the output from the pair
// (2i, 2i + 1).
template <int KIND, bool A16>
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_layer(const uint32_t* num, const uint32_t* den, uint32_t* out_num,
                                                          uint32_t* out_den, uint64_t h) {
  const uint64_t i = (uint64_t)blockIdx.x * GKR_THREADS + threadIdx.x;
  if (i >= h) return;
  if constexpr (KIND == GKR_PROD) {
    stqe<A16>(out_num, i, qm_mul_lazy(ldqe<A16>(num, 2 * i), ldqe<A16>(num, 2 * i + 1)));
  } else {
    const qm31 d0 = ldqe<A16>(den, 2 * i), d1 = ldqe<A16>(den, 2 * i + 1);
    const qm31 n0 = ld_num<KIND, A16>(num, 2 * i), n1 = ld_num<KIND, A16>(num, 2 * i + 1);
    stqe<A16>(out_num, i, frac_num<KIND>(n0, d0, n1, d1));
    stqe<A16>(out_den, i, qm_mul_lazy(d0, d1));
  }
}

// The round polynomial's two sums (cuda/gkr.cu:260-540): for term i the
// pair at t = 0 is (2i, 2i + 1), at t = 1 (2(n + i), 2(n + i) + 1), and at
// t = 2 each element is 2 (t = 1) - (t = 0); F is the pair's product or its
// fraction sum num + lambda den; sum0 += eq[i] F(t = 0), sum2 += eq[i] F(t = 2).
// (A LogUp term is 751 VALU instructions for 144 bytes, so the loads and
// the arithmetic of one grid round take about as long as each other; with
// every wave loading, then computing, the kernel reaches about half of
// either peak. Loading the next term into a second register set while
// computing this one was rescheduled by the compiler into the same
// wait-then-compute order, so it was not kept.)
struct GkrTerm {
  qm31 e, x00, x01, x10, x11, d00, d01, d10, d11;
};
template <int KIND, bool A16>
__device__ __forceinline__ void gkr_load(const uint32_t* eq, const uint32_t* num, const uint32_t* den, uint64_t n,
                                         uint64_t i, GkrTerm& g) {
  const uint64_t a = 2 * i, b = 2 * (n + i);
  g.e = ldqe<A16>(eq, i);
  if constexpr (KIND != GKR_SINGLES) {
    g.x00 = ld_num<KIND, A16>(num, a);
    g.x01 = ld_num<KIND, A16>(num, a + 1);
    g.x10 = ld_num<KIND, A16>(num, b);
    g.x11 = ld_num<KIND, A16>(num, b + 1);
  }
  if constexpr (KIND != GKR_PROD) {
    g.d00 = ldqe<A16>(den, a);
    g.d01 = ldqe<A16>(den, a + 1);
    g.d10 = ldqe<A16>(den, b);
    g.d11 = ldqe<A16>(den, b + 1);
  }
}
template <int KIND>
__device__ __forceinline__ void gkr_add_term(const GkrTerm& g, const qm31& lam, uint64_t (&s0)[4], uint64_t (&s2)[4]) {
  qm31 f0, f2;
  auto twice_minus = [](const qm31& x1, const qm31& x0) { return qm_sub(qm_add(x1, x1), x0); };
  if constexpr (KIND == GKR_PROD) {
    f0 = qm_mul_lazy(g.x00, g.x01);
    f2 = qm_mul_lazy(twice_minus(g.x10, g.x00), twice_minus(g.x11, g.x01));
  } else {
    const qm31 d20 = twice_minus(g.d10, g.d00), d21 = twice_minus(g.d11, g.d01);
    qm31 n00, n01, n20, n21;
    if constexpr (KIND == GKR_SINGLES) {
      n00 = n01 = n20 = n21 = qm31{{1u, 0u, 0u, 0u}};  // 2 * 1 - 1 = 1
    } else {
      n00 = g.x00;
      n01 = g.x01;
      n20 = twice_minus(g.x10, g.x00);
      n21 = twice_minus(g.x11, g.x01);
    }
    f0 = qm_add(frac_num<KIND>(n00, g.d00, n01, g.d01), qm_mul_lazy(lam, qm_mul_lazy(g.d00, g.d01)));
    f2 = qm_add(frac_num<KIND>(n20, d20, n21, d21), qm_mul_lazy(lam, qm_mul_lazy(d20, d21)));
  }
  const qm31 t0 = qm_mul_lazy(g.e, f0), t2 = qm_mul_lazy(g.e, f2);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    s0[c] += t0.v[c];
    s2[c] += t2.v[c];
  }
}

__device__ __forceinline__ uint32_t wave_sum_m31(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = m31_add(v, (uint32_t)__shfl_xor((int)v, off, 64));
  return v;
}

// Grid-stride sums; workgroup b writes its (sum0, sum2) to partial[8b..8b+8).
template <int KIND, bool A16>
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_sum(const uint32_t* eq, const uint32_t* num, const uint32_t* den,
                                                        const uint32_t* lam_p, uint32_t* partial, uint64_t n) {
  __shared__ uint32_t red[GKR_THREADS / 64][8];
  const qm31 lam = KIND == GKR_PROD ? qm31{{0u, 0u, 0u, 0u}} : ldq_arg(lam_p);
  uint64_t s0[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
  const uint64_t stride = (uint64_t)gridDim.x * GKR_THREADS;
  uint64_t i = (uint64_t)blockIdx.x * GKR_THREADS + threadIdx.x;
  for (; i < n; i += stride) {
    GkrTerm g;
    gkr_load<KIND, A16>(eq, num, den, n, i, g);
    gkr_add_term<KIND>(g, lam, s0, s2);
  }
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const uint32_t a = wave_sum_m31(m31_reduce64(s0[c])), b = wave_sum_m31(m31_reduce64(s2[c]));
    if (l == 0) {
      red[w][c] = a;
      red[w][4 + c] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    uint32_t acc = 0;
    for (uint32_t k = 0; k < GKR_THREADS / 64; k++) acc = m31_add(acc, red[k][threadIdx.x]);
    partial[8ull * blockIdx.x + threadIdx.x] = acc;
  }
}

// partial[0..8) = the sum of blocks 0..n_blocks (one workgroup; every block
// is read before block 0 is written).
__global__ __launch_bounds__(GKR_THREADS) void k_gkr_sum_blocks(uint32_t* partial, uint32_t n_blocks) {
  __shared__ uint32_t red[GKR_THREADS / 64][8];
  uint64_t s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < n_blocks; b += GKR_THREADS)
#pragma unroll
    for (int c = 0; c < 8; c++) s[c] += partial[8ull * b + c];
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const uint32_t a = wave_sum_m31(m31_reduce64(s[c]));
    if (l == 0) red[w][c] = a;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    uint32_t acc = 0;
    for (uint32_t k = 0; k < GKR_THREADS / 64; k++) acc = m31_add(acc, red[k][threadIdx.x]);
    partial[threadIdx.x] = acc;
  }
}

bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Workgroups of kernel K resident on the whole device at once (occupancy x
// CUs), once per device: a grid-stride launch of exactly that many runs in
// one round, with no partly filled second round.
template <typename K>
uint32_t resident_wgs(K kernel) {
  static std::atomic<uint32_t> cache[64];
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < 64, "gkr: device index out of range");
  uint32_t n = cache[dev].load(std::memory_order_relaxed);
  if (!n) {
    int per_cu = 0, cus = 0;
    VX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), GKR_THREADS,
                                                        0));
    VX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    n = (uint32_t)((per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1));
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}
dim3 grid_for(uint64_t n) { return dim3((uint32_t)((n + GKR_THREADS - 1) / GKR_THREADS)); }

}  // namespace

void gkr_fix_first_variable(const uint32_t* in, uint32_t* out, const uint32_t* r, uint64_t n, bool secure,
                            hipStream_t s) {
  require(in && out && r, "gkr_fix_first_variable: null pointer");
  require(n >= 2 && (n & 1) == 0 && n / 2 <= (1ull << 31), "gkr_fix_first_variable: n must be even, 2..2^32");
  const uint64_t h = n / 2;
  const bool a = al16(out) && (!secure || al16(in));
  if (secure) {
    if (a) hipLaunchKernelGGL(k_gkr_fix_secure<true>, grid_for(h), dim3(GKR_THREADS), 0, s, in, out, r, h);
    else hipLaunchKernelGGL(k_gkr_fix_secure<false>, grid_for(h), dim3(GKR_THREADS), 0, s, in, out, r, h);
  } else {
    if (a) hipLaunchKernelGGL(k_gkr_fix_base<true>, grid_for(h), dim3(GKR_THREADS), 0, s, in, out, r, h);
    else hipLaunchKernelGGL(k_gkr_fix_base<false>, grid_for(h), dim3(GKR_THREADS), 0, s, in, out, r, h);
  }
  VX_HIP(hipGetLastError());
}

void gkr_gen_eq_evals_init(uint32_t* buf, const uint32_t* v, hipStream_t s) {
  require(buf && v, "gkr_gen_eq_evals_init: null pointer");
  hipLaunchKernelGGL(k_gkr_eq_init, dim3(1), dim3(64), 0, s, buf, v);
  VX_HIP(hipGetLastError());
}

void gkr_gen_eq_evals_pass(uint32_t* buf, const uint32_t* y, uint64_t cur, hipStream_t s) {
  require(buf && y, "gkr_gen_eq_evals_pass: null pointer");
  require(cur >= 1 && cur <= (1ull << 31), "gkr_gen_eq_evals_pass: cur_size must be 1..2^31");
  if (al16(buf)) hipLaunchKernelGGL(k_gkr_eq_pass<true>, grid_for(cur), dim3(GKR_THREADS), 0, s, buf, y, cur);
  else hipLaunchKernelGGL(k_gkr_eq_pass<false>, grid_for(cur), dim3(GKR_THREADS), 0, s, buf, y, cur);
  VX_HIP(hipGetLastError());
}

void gkr_gen_eq_evals(uint32_t* buf, const uint32_t* y, uint32_t k, const uint32_t* v, hipStream_t s) {
  require(buf && v && (k == 0 || y), "gkr_gen_eq_evals: null pointer");
  require(k <= 31, "gkr_gen_eq_evals: at most 31 variables");
  const uint32_t kl = k < GKR_EQ_LO_MAX ? k : GKR_EQ_LO_MAX;
  const uint64_t n_hi = 1ull << (k - kl);
  const uint32_t hpw = (uint32_t)(n_hi < GKR_EQ_HPW ? n_hi : GKR_EQ_HPW);
  hipLaunchKernelGGL(k_gkr_eq_table, dim3((uint32_t)(n_hi / hpw)), dim3(GKR_THREADS), 0, s, buf, y, k, kl, hpw, v,
                     al16(buf));
  VX_HIP(hipGetLastError());
}

void gkr_next_layer(int kind, const uint32_t* num, const uint32_t* den, uint32_t* out_num, uint32_t* out_den,
                    uint64_t n, hipStream_t s) {
  require(kind >= GKR_PROD && kind <= GKR_SINGLES, "gkr_next_layer: unknown layer kind");
  require(n >= 2 && (n & 1) == 0 && n / 2 <= (1ull << 31), "gkr_next_layer: n must be even, 2..2^32");
  require(out_num && (kind == GKR_PROD ? num != nullptr : (den && out_den && (kind == GKR_SINGLES || num))),
          "gkr_next_layer: null pointer");
  const uint64_t h = n / 2;
  const bool a = al16(out_num) && (kind == GKR_PROD || (al16(den) && al16(out_den))) &&
                 (kind != GKR_PROD && kind != GKR_GENERIC ? true : al16(num));
#define VX_GKR_LAYER(K)                                                                                          \
  if (a) hipLaunchKernelGGL((k_gkr_layer<K, true>), grid_for(h), dim3(GKR_THREADS), 0, s, num, den, out_num,     \
                            out_den, h);                                                                         \
  else hipLaunchKernelGGL((k_gkr_layer<K, false>), grid_for(h), dim3(GKR_THREADS), 0, s, num, den, out_num,      \
                          out_den, h);
  switch (kind) {
    case GKR_PROD: VX_GKR_LAYER(GKR_PROD) break;
    case GKR_GENERIC: VX_GKR_LAYER(GKR_GENERIC) break;
    case GKR_MULT: VX_GKR_LAYER(GKR_MULT) break;
    default: VX_GKR_LAYER(GKR_SINGLES) break;
  }
#undef VX_GKR_LAYER
  VX_HIP(hipGetLastError());
}

uint32_t gkr_sum_poly(int kind, const uint32_t* eq, const uint32_t* num, const uint32_t* den, const uint32_t* lambda,
                      uint32_t* partial, uint64_t n_terms, hipStream_t s) {
  require(kind >= GKR_PROD && kind <= GKR_SINGLES, "gkr_sum_poly: unknown layer kind");
  require(n_terms >= 1 && n_terms <= (1ull << 31), "gkr_sum_poly: n_terms must be 1..2^31");
  require(eq && partial && (kind == GKR_PROD ? num != nullptr : (den && lambda && (kind == GKR_SINGLES || num))),
          "gkr_sum_poly: null pointer");
  // at most the ceil(n_terms / 256) blocks the caller's partial buffer
  // holds, and at most what the device holds at once
  const uint64_t need = (n_terms + GKR_THREADS - 1) / GKR_THREADS;
  const bool a = al16(eq) && (kind == GKR_PROD || al16(den)) && (kind != GKR_PROD && kind != GKR_GENERIC ? true : al16(num));
  uint32_t wgs = 1;
#define VX_GKR_SUM(K)                                                                                            \
  if (a) {                                                                                                       \
    const uint64_t r = resident_wgs(k_gkr_sum<K, true>);                                                         \
    wgs = (uint32_t)(need < r ? need : r);                                                                       \
    hipLaunchKernelGGL((k_gkr_sum<K, true>), dim3(wgs), dim3(GKR_THREADS), 0, s, eq, num, den, lambda, partial,    \
                       n_terms);                                                                                 \
  } else {                                                                                                       \
    const uint64_t r = resident_wgs(k_gkr_sum<K, false>);                                                        \
    wgs = (uint32_t)(need < r ? need : r);                                                                       \
    hipLaunchKernelGGL((k_gkr_sum<K, false>), dim3(wgs), dim3(GKR_THREADS), 0, s, eq, num, den, lambda, partial,   \
                       n_terms);                                                                                 \
  }
  switch (kind) {
    case GKR_PROD: VX_GKR_SUM(GKR_PROD) break;
    case GKR_GENERIC: VX_GKR_SUM(GKR_GENERIC) break;
    case GKR_MULT: VX_GKR_SUM(GKR_MULT) break;
    default: VX_GKR_SUM(GKR_SINGLES) break;
  }
#undef VX_GKR_SUM
  if (wgs > 1) hipLaunchKernelGGL(k_gkr_sum_blocks, dim3(1), dim3(GKR_THREADS), 0, s, partial, wgs);
  VX_HIP(hipGetLastError());
  return 1;
}

}  // namespace vx
