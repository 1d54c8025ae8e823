// Twiddle construction for the circle NTT and FRI folds.
//
// Reference behaviour: cuda/twiddles.cu:23-52 (fold twiddle sources),
// :80-105 (coset points), :127-142 (extract_and_squash), and
// cuda/batch_inverse.cu:8-41 (Montgomery batch inverse with inverse(0) = 0).
// All values are exact, so the chunking below gives identical results.
#include "vx_internal.h"

namespace vx {

// p_i = init * step^i. Each thread squares-and-multiplies once for its first
// point and then steps by `step` for the rest of a 16-point run.
static constexpr uint32_t PTS_PER_THREAD = 16;

__global__ void k_coset_points(cpt init, cpt step, uint32_t* xs, uint32_t* ys, uint64_t n) {
  uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * PTS_PER_THREAD;
  if (first >= n) return;
  cpt p = c_mul(init, c_pow(step, first));
  uint64_t end = first + PTS_PER_THREAD < n ? first + PTS_PER_THREAD : n;
  for (uint64_t i = first; i < end; i++) {
    if (xs) xs[i] = p.x;
    if (ys) ys[i] = p.y;
    p = c_mul(p, step);
  }
}
// n a multiple of 16, G = n / 16 lanes: lane g writes points g + m G,
// m < 16, stepping by step^G, so for each m the lanes of a wave write
// consecutive words (the kernel above writes runs of PTS_PER_THREAD words
// per lane, PTS_PER_THREAD words apart).
__global__ void k_coset_points16(cpt init, cpt step, cpt step_g, uint32_t* xs, uint32_t* ys, uint64_t g_lanes) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= g_lanes) return;
  cpt p = c_mul(init, c_pow(step, g));
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) {
    const uint64_t i = g + m * g_lanes;
    if (xs) xs[i] = p.x;
    if (ys) ys[i] = p.y;
    p = c_mul(p, step_g);
  }
}
void coset_points(cpt init, cpt step, uint32_t* xs, uint32_t* ys, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  if (n % 16 == 0) {
    const uint64_t g_lanes = n / 16;
    hipLaunchKernelGGL(k_coset_points16, dim3((uint32_t)((g_lanes + 255) / 256)), dim3(256), 0, s, init, step,
                       c_pow(step, g_lanes), xs, ys, g_lanes);
  } else {
    uint64_t threads = (n + PTS_PER_THREAD - 1) / PTS_PER_THREAD;
    hipLaunchKernelGGL(k_coset_points, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, init, step, xs,
                       ys, n);
  }
  VX_HIP(hipGetLastError());
}

// tw[i] = x[2i]; sq[i] = 2 x[2i]^2 - 1.
__global__ void k_extract_and_squash(const uint32_t* in, uint32_t* tw, uint32_t* sq, uint64_t half_n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  uint32_t x = in[2 * i];
  if (tw) tw[i] = x;
  uint32_t x2 = m31_mul(x, x);
  sq[i] = m31_sub(m31_add(x2, x2), 1u);
}
void extract_and_squash(const uint32_t* in, uint32_t* tw, uint32_t* sq, uint64_t half_n, hipStream_t s) {
  if (half_n == 0) return;
  hipLaunchKernelGGL(k_extract_and_squash, dim3((uint32_t)((half_n + 255) / 256)), dim3(256), 0, s, in, tw, sq,
                     half_n);
  VX_HIP(hipGetLastError());
}

// Batch inverse, 16 elements per lane: prefix products over the non-zero
// entries, one Fermat inversion, then unwind. Zeros map to zero. With
// G = ceil(n / 16) lanes, lane g takes elements g + m G (m < 16), so for
// each m the lanes of a wave read and write consecutive words.
static constexpr uint32_t INV_CHUNK = 16;
__global__ void k_batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t g_lanes) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= g_lanes) return;
  uint32_t v[INV_CHUNK], pre[INV_CHUNK];
  uint32_t acc = 1;
#pragma unroll
  for (uint32_t i = 0; i < INV_CHUNK; i++) {
    const uint64_t e = g + i * g_lanes;
    v[i] = e < n ? in[e] : 0u;
    pre[i] = acc;  // product of the non-zero entries before i
    if (v[i] != 0) acc = m31_mul(acc, v[i]);
  }
  uint32_t inv = m31_inv(acc);  // acc >= 1 is never zero
#pragma unroll
  for (int i = INV_CHUNK - 1; i >= 0; i--) {
    uint32_t r = 0;
    if (v[i] != 0) {
      r = m31_mul(inv, pre[i]);
      inv = m31_mul(inv, v[i]);
    }
    const uint64_t e = g + (uint64_t)i * g_lanes;
    if (e < n) out[e] = r;
  }
}
void batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  const uint64_t g_lanes = (n + INV_CHUNK - 1) / INV_CHUNK;
  hipLaunchKernelGGL(k_batch_inverse, dim3((uint32_t)((g_lanes + 255) / 256)), dim3(256), 0, s, in, out, n,
                     g_lanes);
  VX_HIP(hipGetLastError());
}

// out[i] = coord(init * step^(bitrev_log_n(2i))) for i < n/2.
__global__ void k_fold_tw_sources(cpt init, cpt step, uint32_t* out, uint64_t half_n, uint32_t log_n, int use_y) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half_n) return;
  uint64_t idx = __brevll(i << 1) >> (64 - log_n);
  cpt p = c_mul(init, c_pow(step, idx));
  out[i] = use_y ? p.y : p.x;
}
// The same for h = log_n - 1 >= 4, 16 words per lane: bitrev_log_n(2i) is
// rev_h(i), and for i = g + m 2^(h-4) (m < 16) that is
// (rev_(h-4)(g) << 4) | rev_4(m): one power of step^16 per lane, then two
// multiplications per word with a 16-entry table of step^k; for each m the
// lanes of a wave write consecutive words.
struct StepTable {
  cpt init, step16;
  cpt s[16];  // step^k
};
__global__ void k_fold_tw_sources16(StepTable a, uint32_t* out, uint32_t h, int use_y) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t hb = h - 4;
  if (g >= (1ull << hb)) return;
  const uint64_t jb = hb ? (__brevll(g) >> (64 - hb)) : 0;
  const cpt base = c_mul(a.init, c_pow(a.step16, jb));
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) {
    const cpt sm = a.s[((m & 1u) << 3) | ((m & 2u) << 1) | ((m & 4u) >> 1) | ((m & 8u) >> 3)];
    out[g + ((uint64_t)m << hb)] = use_y ? m31_add(m31_mul(base.x, sm.y), m31_mul(base.y, sm.x))
                                         : m31_sub(m31_mul(base.x, sm.x), m31_mul(base.y, sm.y));
  }
}

static StepTable step_table(cpt init, cpt step) {
  StepTable a{};
  a.init = init;
  a.s[0] = c_identity();
  for (int k = 1; k < 16; k++) a.s[k] = c_mul(a.s[k - 1], step);
  a.step16 = c_mul(a.s[15], step);
  return a;
}

void fold_twiddle_sources(cpt init, cpt step, uint32_t* out, uint64_t n, uint32_t log_n, bool extract_y,
                          hipStream_t s) {
  require(log_n >= 1 && log_n <= 31 && n == (1ull << log_n), "fold_twiddle_sources: n must be 2^log_n");
  uint64_t half = n / 2;
  if (log_n >= 5) {
    const uint64_t lanes = 1ull << (log_n - 5);
    hipLaunchKernelGGL(k_fold_tw_sources16, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, s,
                       step_table(init, step), out, log_n - 1, extract_y ? 1 : 0);
  } else {
    hipLaunchKernelGGL(k_fold_tw_sources, dim3((uint32_t)((half + 255) / 256)), dim3(256), 0, s, init, step, out,
                       half, log_n, extract_y ? 1 : 0);
  }
  VX_HIP(hipGetLastError());
}

// stwo flat twiddle buffer of a 2^log_n-point transform (src/ntt.rs:296-354):
// level l is the half_odds(log_n - 1) coset doubled l times; its 2^b words
// (b = log_n - 2 - l) sit at [n/2 - 2^(b+1), n/2 - 2^b) and hold, at position
// bitrev_b(j), the x of that coset's point j, which is pi^l(x(init * step^j))
// with pi(x) = 2x^2 - 1 (doubling a point squares it). The last word is 1.
__global__ void k_stwo_twiddles(cpt init, cpt step, uint32_t* out, uint32_t log_n, uint64_t i_from) {
  const uint64_t half = 1ull << (log_n - 1);
  const uint64_t i = i_from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half) return;
  if (i == half - 1) {
    out[i] = 1u;
    return;
  }
  const uint64_t r = half - 1 - i;          // in [2^b, 2^(b+1)) for level size 2^b
  const uint32_t b = 63u - (uint32_t)__clzll(r);
  const uint64_t pos = i + (2ull << b) - half;
  const uint64_t j = b == 0 ? 0 : (__brevll(pos) >> (64 - b));
  uint32_t x = c_mul(init, c_pow(step, j)).x;
  for (uint32_t l = log_n - 2 - b; l > 0; l--) {
    const uint32_t x2 = m31_mul(x, x);
    x = m31_sub(m31_add(x2, x2), 1u);
  }
  out[i] = x;
}
// Levels of 2^b >= 16 words, 16 words per lane: lane g of level b writes
// positions pos = g + m 2^(b-4), m < 16, whose bit reverses are
// (rev_(b-4)(g) << 4) | rev_4(m). So the lane forms one point
// init * (step^16)^rev(g) and takes each word's x from it and step^rev_4(m)
// (a 16-entry table) with two multiplications, instead of a power of step
// per word; for each m the lanes of a wave write consecutive words. Lane
// u + 1 in [2^(b-4), 2^(b-3)) belongs to level b.
struct StwoTwArgs {
  cpt init, step16;
  cpt s[16];  // step^k
};
__global__ void k_stwo_twiddles_levels(StwoTwArgs a, uint32_t* out, uint32_t log_n) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (u >= (1ull << (log_n - 5))) return;
  const uint32_t b4 = 63u - (uint32_t)__clzll(u);  // b - 4
  const uint64_t g = u - (1ull << b4);
  const uint64_t jb = b4 ? (__brevll(g) >> (64 - b4)) : 0;
  const cpt base = c_mul(a.init, c_pow(a.step16, jb));
  const uint64_t half = 1ull << (log_n - 1);
  uint32_t* o = out + (half - (2ull << (b4 + 4))) + g;
  const uint32_t l = log_n - 6 - b4;  // doublings: log_n - 2 - b
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) {
    const cpt sm = a.s[((m & 1u) << 3) | ((m & 2u) << 1) | ((m & 4u) >> 1) | ((m & 8u) >> 3)];
    uint32_t x = m31_sub(m31_mul(base.x, sm.x), m31_mul(base.y, sm.y));
    for (uint32_t k = 0; k < l; k++) {
      const uint32_t x2 = m31_mul(x, x);
      x = m31_sub(m31_add(x2, x2), 1u);
    }
    o[(uint64_t)m << b4] = x;
  }
}

void stwo_twiddles(uint32_t log_n, uint32_t* tw, uint32_t* itw, hipStream_t s) {
  require(log_n >= 2 && log_n <= 31, "stwo_twiddles: log_n must be 2..31");
  const coset c = make_coset(HALF_ODDS, log_n - 1);
  const uint64_t half = 1ull << (log_n - 1);
  // levels 0..3 and the pad word (the last 16 words, or all of a small
  // buffer) one word per lane
  const uint64_t from = log_n >= 6 ? half - 16 : 0;
  hipLaunchKernelGGL(k_stwo_twiddles, dim3((uint32_t)((half - from + 255) / 256)), dim3(256), 0, s, c.init, c.step,
                     tw, log_n, from);
  if (log_n >= 6) {
    StwoTwArgs a{};
    a.init = c.init;
    a.s[0] = c_identity();
    for (int k = 1; k < 16; k++) a.s[k] = c_mul(a.s[k - 1], c.step);
    a.step16 = c_mul(a.s[15], c.step);
    const uint64_t lanes = (1ull << (log_n - 5)) - 1;
    hipLaunchKernelGGL(k_stwo_twiddles_levels, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, s, a, tw,
                       log_n);
  }
  VX_HIP(hipGetLastError());
  if (itw) batch_inverse(tw, itw, half, s);  // entries are x of points of order >= 8: never 0
}

}  // namespace vx
