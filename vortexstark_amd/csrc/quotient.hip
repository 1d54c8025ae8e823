// The stwo plugin's QuotientOps on the device (crates/cuda-backend/src/
// quotient_ops.rs:119-342).
//
// accumulate_numerators (cuda/quotient.cu:13-50): for each row,
//   out[row] = sum_i (c_i * col[idx_i][row] - b_i)          (QM31, SoA out)
// Here a lane takes 4 rows (256 apart, coalesced) and the columns 4 at a
// time: the column pointers and the c_i are wave-uniform scalars, the 16
// column values are independent loads in flight together, and the M31
// products are summed four at a time in 64 bits before a reduction;
// sum_i b_i is formed once per lane. The fused form
// (cuda/quotient.cu:56-131) computes the z and z_next sums from one read of
// each column value.
//
// compute_quotients_combine (cuda/quotient.cu:153-233): for each row with
// domain point d and every accumulator j (sample point p_j),
//   q[row] = sum_j (numer_j[lifted(row)] - a_j * d.y) / den_j(d),
//   den_j(d) = (Re p.x - d.x) Im p.y - (Re p.y - d.y) Im p.x    (CM31),
// dividing each CM31 half of the QM31 numerator. The reference inverts every
// den_j(d) with its own Fermat power (about 60 multiplications). Here a lane
// takes 8 rows (coalesced, 256 apart) and inverts the 8 norms of each
// accumulator with one Fermat power (Montgomery's trick; a zero norm maps to
// zero, as the reference's inverse does), so the inversions cost an eighth.
// lifted(row) = ((row >> (lift - log_size_j + 1)) << 1) + (row & 1).
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t QN_THREADS = 256;
static constexpr uint32_t QN_ROWS = 4;   // rows per lane, QN_THREADS apart
static constexpr uint32_t QN_GROUP = 4;  // columns per step: 4 products per 64-bit sum

// Sums for one side (z or z_next): 64-bit lazy sums per row and word, the
// reduced running sums, and sum_i b_i.
struct QnAcc {
  uint64_t a[QN_ROWS][4];
  uint32_t r[QN_ROWS][4];
  uint32_t b[4];
};

__device__ __forceinline__ void qn_reduce(QnAcc& s) {
#pragma unroll
  for (uint32_t rr = 0; rr < QN_ROWS; rr++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      s.r[rr][k] = m31_add(s.r[rr][k], m31_reduce64(s.a[rr][k]));
      s.a[rr][k] = 0;
    }
}

// Each step takes QN_GROUP columns: their indices, pointers and
// coefficients are wave-uniform scalar loads issued together, then the
// QN_GROUP x QN_ROWS column values are independent loads in flight at once.
// VEC: a lane takes 4 consecutive rows instead (n_rows a multiple of 4,
// outputs 16-byte aligned): one 16-byte load per column when the column is
// aligned (a wave-uniform check), one 16-byte store per output word.
template <bool DUAL, bool VEC>
__global__ __launch_bounds__(QN_THREADS) void k_accumulate_numerators(
    const uint32_t* const* cols, const uint32_t* idx_z, const uint32_t* zn_idx, const uint32_t* b_z,
    const uint32_t* c_z, const uint32_t* b_zn, const uint32_t* c_zn, uint32_t n_cols, uint32_t n_rows,
    uint32_t* oz0, uint32_t* oz1, uint32_t* oz2, uint32_t* oz3, uint32_t* on0, uint32_t* on1, uint32_t* on2,
    uint32_t* on3) {
  const uint32_t row0 = VEC ? QN_ROWS * (blockIdx.x * QN_THREADS + threadIdx.x)
                           : blockIdx.x * (QN_THREADS * QN_ROWS) + threadIdx.x;
  const uint32_t rstep = VEC ? 1u : QN_THREADS;  // rows of a lane: row0 + rr * rstep
  if (VEC && row0 >= n_rows) return;
  QnAcc z{}, n{};
  for (uint32_t i0 = 0; i0 < n_cols; i0 += QN_GROUP) {
    const uint32_t cnt = n_cols - i0 < QN_GROUP ? n_cols - i0 : QN_GROUP;
    const uint32_t* col[QN_GROUP];
    qm31 cz[QN_GROUP], cn[QN_GROUP];
    bool has_n[QN_GROUP];
#pragma unroll
    for (uint32_t g = 0; g < QN_GROUP; g++) {
      col[g] = nullptr;
      has_n[g] = false;
      if (g < cnt) {
        col[g] = ld_uniform(cols + ld_uniform(idx_z + i0 + g));
        cz[g] = ld_uniform(reinterpret_cast<const qm31*>(c_z) + i0 + g);
        const qm31 bz = ld_uniform(reinterpret_cast<const qm31*>(b_z) + i0 + g);
#pragma unroll
        for (int k = 0; k < 4; k++) z.b[k] = m31_add(z.b[k], bz.v[k]);
        if constexpr (DUAL) {
          const uint32_t zi = ld_uniform(zn_idx + i0 + g);
          if (zi != 0xFFFFFFFFu) {
            has_n[g] = true;
            cn[g] = ld_uniform(reinterpret_cast<const qm31*>(c_zn) + zi);
            const qm31 bn = ld_uniform(reinterpret_cast<const qm31*>(b_zn) + zi);
#pragma unroll
            for (int k = 0; k < 4; k++) n.b[k] = m31_add(n.b[k], bn.v[k]);
          }
        }
      }
    }
    uint32_t f[QN_GROUP][QN_ROWS];
#pragma unroll
    for (uint32_t g = 0; g < QN_GROUP; g++) {
      if (VEC && g < cnt && (reinterpret_cast<uintptr_t>(col[g]) & 15u) == 0) {
        const uint4 v = *reinterpret_cast<const uint4*>(col[g] + row0);
        f[g][0] = v.x;
        f[g][1] = v.y;
        f[g][2] = v.z;
        f[g][3] = v.w;
        continue;
      }
#pragma unroll
      for (uint32_t rr = 0; rr < QN_ROWS; rr++) {
        const uint32_t row = row0 + rr * rstep;
        f[g][rr] = (g < cnt && row < n_rows) ? col[g][row] : 0u;
      }
    }
#pragma unroll
    for (uint32_t g = 0; g < QN_GROUP; g++) {
      if (g >= cnt) break;
#pragma unroll
      for (uint32_t rr = 0; rr < QN_ROWS; rr++)
#pragma unroll
        for (int k = 0; k < 4; k++) z.a[rr][k] += (uint64_t)cz[g].v[k] * f[g][rr];
      if constexpr (DUAL) {
        if (has_n[g]) {
#pragma unroll
          for (uint32_t rr = 0; rr < QN_ROWS; rr++)
#pragma unroll
            for (int k = 0; k < 4; k++) n.a[rr][k] += (uint64_t)cn[g].v[k] * f[g][rr];
        }
      }
    }
    qn_reduce(z);  // at most QN_GROUP products per 64-bit sum: < 4 P^2 < 2^64
    if constexpr (DUAL) qn_reduce(n);
  }
  uint32_t* const oz[4] = {oz0, oz1, oz2, oz3};
  uint32_t* const on[4] = {on0, on1, on2, on3};
  if constexpr (VEC) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      *reinterpret_cast<uint4*>(oz[k] + row0) =
          make_uint4(m31_sub(z.r[0][k], z.b[k]), m31_sub(z.r[1][k], z.b[k]), m31_sub(z.r[2][k], z.b[k]),
                     m31_sub(z.r[3][k], z.b[k]));
      if constexpr (DUAL)
        *reinterpret_cast<uint4*>(on[k] + row0) =
            make_uint4(m31_sub(n.r[0][k], n.b[k]), m31_sub(n.r[1][k], n.b[k]), m31_sub(n.r[2][k], n.b[k]),
                       m31_sub(n.r[3][k], n.b[k]));
    }
    return;
  }
#pragma unroll
  for (uint32_t rr = 0; rr < QN_ROWS; rr++) {
    const uint32_t row = row0 + rr * QN_THREADS;
    if (row >= n_rows) continue;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      oz[k][row] = m31_sub(z.r[rr][k], z.b[k]);
      if constexpr (DUAL) on[k][row] = m31_sub(n.r[rr][k], n.b[k]);
    }
  }
}

void accumulate_numerators(const uint32_t* const* cols, const uint32_t* idx_z, const uint32_t* zn_idx,
                           const uint32_t* b_z, const uint32_t* c_z, const uint32_t* b_zn, const uint32_t* c_zn,
                           uint32_t n_cols, uint32_t n_rows, bool dual, uint32_t* const oz[4],
                           uint32_t* const on[4], hipStream_t s) {
  if (n_rows == 0) return;
  const uint32_t per_block = QN_THREADS * QN_ROWS;
  const dim3 grid((n_rows + per_block - 1) / per_block);
  uintptr_t bits = 0;
  for (int k = 0; k < 4; k++) bits |= reinterpret_cast<uintptr_t>(oz[k]) | (dual ? reinterpret_cast<uintptr_t>(on[k]) : 0);
  const bool vec = n_rows % QN_ROWS == 0 && (bits & 15u) == 0;
  uint32_t* const* o2 = dual ? on : oz;
#define VX_QN_LAUNCH(D, V)                                                                                     \
  hipLaunchKernelGGL((k_accumulate_numerators<D, V>), grid, dim3(QN_THREADS), 0, s, cols, idx_z, zn_idx, b_z, \
                     c_z, b_zn, c_zn, n_cols, n_rows, oz[0], oz[1], oz[2], oz[3], o2[0], o2[1], o2[2], o2[3])
  if (dual && vec) VX_QN_LAUNCH(true, true);
  else if (dual) VX_QN_LAUNCH(true, false);
  else if (vec) VX_QN_LAUNCH(false, true);
  else VX_QN_LAUNCH(false, false);
#undef VX_QN_LAUNCH
  VX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
static constexpr uint32_t QC_ROWS = 8;  // rows per lane, QC_THREADS apart
static constexpr uint32_t QC_THREADS = 256;

struct QcArgs {
  const uint32_t* sp_x;
  const uint32_t* sp_y;
  const uint32_t* fla;
  const uint32_t* const* numer[4];
  const uint32_t* acc_log_sizes;
  uint32_t n_accs;
  const uint32_t* xs;
  const uint32_t* ys;
  uint32_t lifting_log_size, n_rows;
  uint32_t* out[4];
};

__global__ __launch_bounds__(QC_THREADS) void k_quotients_combine(QcArgs a) {
  const uint32_t base = blockIdx.x * (QC_THREADS * QC_ROWS) + threadIdx.x;
  uint32_t dx[QC_ROWS], dy[QC_ROWS];
  qm31 q[QC_ROWS];
#pragma unroll
  for (uint32_t r = 0; r < QC_ROWS; r++) {
    const uint32_t row = base + r * QC_THREADS;
    const bool in = row < a.n_rows;
    dx[r] = in ? a.xs[row] : 0u;
    dy[r] = in ? a.ys[row] : 0u;
    q[r] = qm31{{0u, 0u, 0u, 0u}};
  }
  for (uint32_t j = 0; j < a.n_accs; j++) {
    const qm31 px = ld_uniform(reinterpret_cast<const qm31*>(a.sp_x) + j);
    const qm31 py = ld_uniform(reinterpret_cast<const qm31*>(a.sp_y) + j);
    const qm31 al = ld_uniform(reinterpret_cast<const qm31*>(a.fla) + j);
    // an accumulator larger than the lifting domain (a caller error) reads
    // its first rows instead of out of bounds; so does one of a single row
    // (log size 0, which the lifting formula does not cover): every index is
    // clamped to the accumulator's last row
    const uint32_t alog = ld_uniform(a.acc_log_sizes + j);
    const uint32_t lr = alog <= a.lifting_log_size ? a.lifting_log_size - alog : 0u;
    const uint32_t sh = lr < 31u ? lr + 1u : 31u;
    const uint32_t last = alog < 31u ? (1u << alog) - 1u : 0x7FFFFFFFu;
    const uint32_t* nu[4];
#pragma unroll
    for (int k = 0; k < 4; k++) nu[k] = ld_uniform(a.numer[k] + j);
    const cm31 prx{px.v[0], px.v[1]}, pix{px.v[2], px.v[3]};
    const cm31 pry{py.v[0], py.v[1]}, piy{py.v[2], py.v[3]};
    cm31 den[QC_ROWS];
    uint32_t pre[QC_ROWS];
    uint32_t acc = 1;
#pragma unroll
    for (uint32_t r = 0; r < QC_ROWS; r++) {
      den[r] = cm_sub(cm_mul(cm_sub(prx, cm31{dx[r], 0u}), piy), cm_mul(cm_sub(pry, cm31{dy[r], 0u}), pix));
      const uint32_t nrm = m31_add(m31_mul(den[r].a, den[r].a), m31_mul(den[r].b, den[r].b));
      pre[r] = acc;  // product of the non-zero norms before r
      if (nrm != 0u) acc = m31_mul(acc, nrm);
    }
    uint32_t inv = m31_inv(acc);
#pragma unroll
    for (int r = QC_ROWS - 1; r >= 0; r--) {
      const uint32_t nrm = m31_add(m31_mul(den[r].a, den[r].a), m31_mul(den[r].b, den[r].b));
      uint32_t inv_n = 0u;
      if (nrm != 0u) {
        inv_n = m31_mul(inv, pre[r]);
        inv = m31_mul(inv, nrm);
      }
      const uint32_t row = base + r * QC_THREADS;
      if (row >= a.n_rows) continue;
      const cm31 di{m31_mul(den[r].a, inv_n), m31_mul(m31_neg(den[r].b), inv_n)};  // conj / norm
      const uint32_t li = min(((row >> sh) << 1) + (row & 1u), last);
      qm31 full;
#pragma unroll
      for (int k = 0; k < 4; k++) full.v[k] = m31_sub(nu[k][li], m31_mul(al.v[k], dy[r]));
      const cm31 ra = cm_mul(cm31{full.v[0], full.v[1]}, di), rb = cm_mul(cm31{full.v[2], full.v[3]}, di);
      q[r] = qm_add(q[r], qm31{{ra.a, ra.b, rb.a, rb.b}});
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < QC_ROWS; r++) {
    const uint32_t row = base + r * QC_THREADS;
    if (row >= a.n_rows) continue;
#pragma unroll
    for (int k = 0; k < 4; k++) a.out[k][row] = q[r].v[k];
  }
}

void quotients_combine(const uint32_t* sp_x, const uint32_t* sp_y, const uint32_t* fla,
                       const uint32_t* const* const numer[4], const uint32_t* acc_log_sizes, uint32_t n_accs,
                       const uint32_t* xs, const uint32_t* ys, uint32_t lifting_log_size, uint32_t n_rows,
                       uint32_t* const out[4], hipStream_t s) {
  if (n_rows == 0) return;
  QcArgs a{sp_x, sp_y, fla, {numer[0], numer[1], numer[2], numer[3]}, acc_log_sizes, n_accs, xs, ys,
           lifting_log_size, n_rows, {out[0], out[1], out[2], out[3]}};
  const uint32_t per_block = QC_THREADS * QC_ROWS;
  hipLaunchKernelGGL(k_quotients_combine, dim3((n_rows + per_block - 1) / per_block), dim3(QC_THREADS), 0, s, a);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
