// Column operations of the stwo plugin (crates/cuda-backend/src/column.rs,
// accumulation_ops.rs): element-wise M31 add / mul, the lifted accumulation
// index gathers and the 16-way leaf packing.
//
// Reference kernels: cuda_m31_add / cuda_m31_mul (src/cuda/ffi.rs:212-213),
// cuda/accumulate_lift.cu:14-40, cuda/gather.cu and the bounds-checked
// cuda/forge/gather_forge.cu:82-116. The element-wise kernels move 16 bytes
// per lane (uint4) when the buffers allow it; lift_and_accumulate fuses the
// plugin's per-column loop into one pass over the largest column.
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t COL_THREADS = 256;

static bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <bool MUL>
__global__ void k_m31_binop4(const uint4* a, const uint4* b, uint4* out, uint64_t n4) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const uint4 x = a[i], y = b[i];
  out[i] = MUL ? make_uint4(m31_mul(x.x, y.x), m31_mul(x.y, y.y), m31_mul(x.z, y.z), m31_mul(x.w, y.w))
               : make_uint4(m31_add(x.x, y.x), m31_add(x.y, y.y), m31_add(x.z, y.z), m31_add(x.w, y.w));
}
template <bool MUL>
__global__ void k_m31_binop(const uint32_t* a, const uint32_t* b, uint32_t* out, uint64_t from, uint64_t n) {
  const uint64_t i = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = MUL ? m31_mul(a[i], b[i]) : m31_add(a[i], b[i]);
}

void m31_binop(const uint32_t* a, const uint32_t* b, uint32_t* out, uint64_t n, bool mul, hipStream_t s) {
  if (n == 0) return;
  uint64_t from = 0;
  if (al16(a) && al16(b) && al16(out) && n >= 4) {
    const uint64_t n4 = n / 4;
    const dim3 g((uint32_t)((n4 + COL_THREADS - 1) / COL_THREADS));
    if (mul)
      hipLaunchKernelGGL(k_m31_binop4<true>, g, dim3(COL_THREADS), 0, s, reinterpret_cast<const uint4*>(a),
                         reinterpret_cast<const uint4*>(b), reinterpret_cast<uint4*>(out), n4);
    else
      hipLaunchKernelGGL(k_m31_binop4<false>, g, dim3(COL_THREADS), 0, s, reinterpret_cast<const uint4*>(a),
                         reinterpret_cast<const uint4*>(b), reinterpret_cast<uint4*>(out), n4);
    from = n4 * 4;
  }
  if (from < n) {
    const dim3 g((uint32_t)((n - from + COL_THREADS - 1) / COL_THREADS));
    if (mul)
      hipLaunchKernelGGL(k_m31_binop<true>, g, dim3(COL_THREADS), 0, s, a, b, out, from, n);
    else
      hipLaunchKernelGGL(k_m31_binop<false>, g, dim3(COL_THREADS), 0, s, a, b, out, from, n);
  }
  VX_HIP(hipGetLastError());
}

// col[i] += curr[((i >> (log_ratio + 1)) << 1) | (i & 1)]
__global__ void k_accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= col_n) return;
  col[i] = m31_add(col[i], curr[((i >> (log_ratio + 1)) << 1) | (i & 1u)]);
}

void accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio, hipStream_t s) {
  if (col_n == 0) return;
  hipLaunchKernelGGL(k_accumulate_lift, dim3((col_n + COL_THREADS - 1) / COL_THREADS), dim3(COL_THREADS), 0, s,
                     col, curr, col_n, log_ratio);
  VX_HIP(hipGetLastError());
}

// The plugin's lift_and_accumulate (accumulation_ops.rs:44-85) applies
// accumulate_lift column by column, smallest first, each onto the running
// sum. Lifting composes: ((j >> (r2 + 1)) << 1) | (j & 1) of
// j = ((i >> (r1 + 1)) << 1) | (i & 1) is the lift of i by r1 + r2. So the
// largest column's final value is
//   last[i] + sum_k col_k[((i >> (L - L_k + 1)) << 1) | (i & 1)],
// computed here in one pass per channel (blockIdx.y) over the largest
// column; the smaller columns are read, not written.
__global__ void k_lift_accumulate(const uint32_t* const* d_src, const uint32_t* d_shift, uint32_t n_src,
                                  uint32_t* const* d_last, uint32_t last_n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t ch = blockIdx.y;
  if (i >= last_n) return;
  uint32_t* last = ld_uniform(d_last + ch);
  uint32_t acc = last[i];
  for (uint32_t k = 0; k < n_src; k++) {
    const uint32_t* c = ld_uniform(d_src + (uint64_t)ch * n_src + k);
    const uint32_t sh = ld_uniform(d_shift + k);
    acc = m31_add(acc, c[((i >> sh) << 1) | (i & 1u)]);
  }
  last[i] = acc;
}

void lift_accumulate(const uint32_t* const* d_src, const uint32_t* d_shift, uint32_t n_src, uint32_t* const* d_last,
                     uint32_t last_n, hipStream_t s) {
  if (last_n == 0) return;
  hipLaunchKernelGGL(k_lift_accumulate, dim3((last_n + COL_THREADS - 1) / COL_THREADS, 4), dim3(COL_THREADS), 0, s,
                     d_src, d_shift, n_src, d_last, last_n);
  VX_HIP(hipGetLastError());
}

// dst[i] = src[idx[i]] (W words per element); with src_len set, elements
// whose source words are not all below src_len are skipped, as the
// bounds-checked reference variant does.
template <uint32_t W>
__global__ void k_gather(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, uint64_t src_len,
                         int checked) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t off = (uint64_t)idx[i] * W;
  if (checked && off + W - 1 >= src_len) return;
#pragma unroll
  for (uint32_t w = 0; w < W; w++) dst[(uint64_t)i * W + w] = src[off + w];
}

void gather(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, uint32_t words, bool checked,
            uint64_t src_len, hipStream_t s) {
  if (n == 0) return;
  const dim3 g((n + COL_THREADS - 1) / COL_THREADS);
  if (words == 1)
    hipLaunchKernelGGL(k_gather<1>, g, dim3(COL_THREADS), 0, s, src, idx, dst, n, src_len, checked ? 1 : 0);
  else
    hipLaunchKernelGGL(k_gather<8>, g, dim3(COL_THREADS), 0, s, src, idx, dst, n, src_len, checked ? 1 : 0);
  VX_HIP(hipGetLastError());
}

// The plugin's PackLeavesOps (merkle_ops.rs:197-228, cuda/pack_leaves.cu:
// 13-35): 4 M31 columns of N words into 64 columns of N / 16 words,
// out[16 c + (p mod 16)][p / 16] = in[c][p]. A lane takes the 16 consecutive
// words of one output row (four 16-byte loads when the column is aligned)
// and writes one word to each of the 16 output columns of its coordinate;
// neighbouring lanes write neighbouring words.
__global__ void k_pack_leaves(const uint32_t* const* in, uint32_t* const* out, uint32_t rows) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;  // output row
  const uint32_t c = blockIdx.y;                             // coordinate
  if (r >= rows) return;
  const uint32_t* src = ld_uniform(in + c) + (uint64_t)r * 16;
  uint32_t v[16];
  if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0) {
#pragma unroll
    for (uint32_t m = 0; m < 4; m++) {
      const uint4 x = reinterpret_cast<const uint4*>(src)[m];
      v[4 * m] = x.x;
      v[4 * m + 1] = x.y;
      v[4 * m + 2] = x.z;
      v[4 * m + 3] = x.w;
    }
  } else {
#pragma unroll
    for (uint32_t m = 0; m < 16; m++) v[m] = src[m];
  }
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) ld_uniform(out + 16 * c + m)[r] = v[m];
}

void pack_leaves(const uint32_t* const* d_in, uint32_t* const* d_out, uint32_t n, hipStream_t s) {
  const uint32_t rows = n / 16;
  if (rows == 0) return;
  hipLaunchKernelGGL(k_pack_leaves, dim3((rows + COL_THREADS - 1) / COL_THREADS, 4), dim3(COL_THREADS), 0, s, d_in,
                     d_out, rows);
  VX_HIP(hipGetLastError());
}

// Half-coset natural order <-> canonic bit-reversed order (cuda/permute.cu:
// 30-57, src/circle.rs:192-220). Natural index k of half_coset(log_n) is
// canonic index cn = k / 2 (k even) or n - 1 - k / 2 (k odd), stored at
// rev(cn). With h = log_n - 1, M = 2^h - 1 and i = 2m + p this is
//   F(i) = ((rev_h(m) ^ (p ? M : 0)) << 1) | p,
// an involution (rev_h(M) = M), so the reference's inverse,
// dst[k] = src[F(k)], is dst[F(j)] = src[j] again: both entry points run the
// one permutation dst[F(i)] = src[i].
__device__ __forceinline__ uint32_t perm_f(uint32_t i, uint32_t h) {
  const uint32_t m = i >> 1, p = i & 1u;
  const uint32_t r = h ? __brev(m) >> (32 - h) : 0u;
  const uint32_t M = h ? 0xFFFFFFFFu >> (32 - h) : 0u;
  return ((r ^ (p ? M : 0u)) << 1) | p;
}

// Small columns (log_n <= 10) and buffers not 16-byte aligned: one element
// per lane, scattered.
__global__ void k_permute_scatter(const uint32_t* src, uint32_t* dst, uint32_t n, uint32_t h) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[perm_f(i, h)] = src[i];
}

// log_n >= 11: i = (a, mid, b, p) with a, b of PM_T bits and mid of
// mb = h - 2 PM_T bits; F(i) = (rev b ^ f, rev mid ^ f', rev a ^ f, p), f
// and f' all ones when p = 1. A run of 64 consecutive destinations (fixed
// B, MIDd) takes its even words from source middle value rev(MIDd) and its
// odd words from the complement of that, so a workgroup loads the pair of
// 32 x 64 tiles {mid, ~mid} (32 runs of 64 consecutive source words each)
// and writes 64 full destination runs of 64 words: 256-byte runs on both
// sides, 16 bytes per lane. With mb = 0 there is one tile, which pairs with
// itself.
static constexpr uint32_t PM_T = 5, PM_W = 1u << PM_T;
__global__ __launch_bounds__(256) void k_permute_tiles(const uint32_t* src, uint32_t* dst, uint32_t h) {
  __shared__ uint32_t tile[2][PM_W][2 * PM_W + 1];
  const uint32_t mb = h - 2 * PM_T;
  const uint32_t t = threadIdx.x, u = t & 15u;  // u: 16-byte column of a 64-word row or run
  const uint32_t mmask = mb ? 0xFFFFFFFFu >> (32 - mb) : 0u;
  const uint32_t mid0 = blockIdx.x, mid1 = mid0 ^ mmask;  // blockIdx.x < 2^(mb - 1): mid0's top bit is 0
  const uint32_t ntiles = mb ? 2u : 1u;
  const uint32_t ash = mb + PM_T + 1;  // position of a (source) and B (destination)
  // 16 lanes per 64-word row, 16 rows per pass, one 16-byte load each
#pragma unroll
  for (uint32_t q = 0; q < 2; q++) {
    if (q >= ntiles) break;
    const uint64_t base = (uint64_t)(q ? mid1 : mid0) << (PM_T + 1);
#pragma unroll
    for (uint32_t r = t >> 4; r < PM_W; r += 16) {
      const uint4 v = *reinterpret_cast<const uint4*>(src + (((uint64_t)r << ash) | base | (4 * u)));
      tile[q][r][4 * u] = v.x;
      tile[q][r][4 * u + 1] = v.y;
      tile[q][r][4 * u + 2] = v.z;
      tile[q][r][4 * u + 3] = v.w;
    }
  }
  __syncthreads();
  const uint32_t rmid0 = mb ? __brev(mid0) >> (32 - mb) : 0u;
  // destination word e = 4u + k of a run is entry (x = e >> 1, p = e & 1)
  uint32_t a[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    const uint32_t e = 4 * u + k, flip = (e & 1u) ? PM_W - 1 : 0u;
    a[k] = __brev((e >> 1) ^ flip) >> (32 - PM_T);
  }
#pragma unroll 4
  for (uint32_t run = t >> 4; run < ntiles * PM_W; run += 16) {
    const uint32_t B = run & (PM_W - 1), which = run >> PM_T;
    const uint32_t midd = which ? rmid0 ^ mmask : rmid0;
    const uint32_t be = __brev(B) >> (32 - PM_T), bo = __brev(B ^ (PM_W - 1)) >> (32 - PM_T);
    const uint32_t te = mb ? which : 0u, to = mb ? which ^ 1u : 0u;  // tiles of the even / odd words
    const uint4 v = make_uint4(tile[te][a[0]][2 * be], tile[to][a[1]][2 * bo + 1], tile[te][a[2]][2 * be],
                               tile[to][a[3]][2 * bo + 1]);
    *reinterpret_cast<uint4*>(dst + (((uint64_t)B << ash) | ((uint64_t)midd << (PM_T + 1)) | (4 * u))) = v;
  }
}

void permute_hc_brt(const uint32_t* src, uint32_t* dst, uint32_t log_n, hipStream_t s) {
  const uint32_t h = log_n ? log_n - 1 : 0;
  const bool aligned = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) == 0;
  if (log_n <= 2 * PM_T || !aligned) {
    const uint32_t n = 1u << log_n;
    hipLaunchKernelGGL(k_permute_scatter, dim3((n + COL_THREADS - 1) / COL_THREADS), dim3(COL_THREADS), 0, s, src,
                       dst, n, h);
  } else {
    const uint32_t mb = h - 2 * PM_T;
    hipLaunchKernelGGL(k_permute_tiles, dim3(mb ? 1u << (mb - 1) : 1u), dim3(256), 0, s, src, dst, h);
  }
  VX_HIP(hipGetLastError());
}

}  // namespace vx
