// Blake2s Merkle commitments for gfx950.
//
// Tree semantics (cuda/blake2s.cu:13-87, src/merkle.rs:520-593,841-884):
// leaf = Blake2s(row words, t = 4 * n_cols), node = Blake2s(left || right,
// t = 64), leaves in natural order, parent j = H(2j, 2j+1).
//
// MI355X design. A 256-lane workgroup commits one tile of leaves: each lane
// hashes 2^LPT consecutive leaves and folds them to one node entirely in
// VGPRs (state and message in registers), then the tile's upper levels are
// merged through LDS. Big trees (>= 2^21 leaves) use 16 leaves per lane and
// 4096-leaf tiles (throughput); smaller ones use 4 leaves per lane and
// 1024-leaf tiles, which cuts a workgroup's dependent chain from 39 to 15
// compressions (latency). Only levels >= 4 are written to HBM (about 4 bytes
// per leaf); the decommitment re-hashes the 16-leaf group around a query.
// Leaf words can be produced on the fly by fused producers, so the Fibonacci
// quotient is never materialised and each FRI fold writes its layer once
// while hashing it:
//   LEAF_QUOTIENT       q_i = alpha (e[i+2] - e[i+1] - e[i])  (cuda/constraint.cu:10-36)
//   LEAF_FOLD_CIRCLE_Q  circle fold of that quotient           (cuda/fri.cu:92-126)
//   LEAF_FOLD_LINE      line fold of the previous layer        (cuda/fri.cu:58-87)
// The loads of the next 4-leaf group are issued before the current group is
// hashed. The kernel that produces a root also runs the transcript step
// (mix the root, draw the next alpha), so no separate channel launch is
// needed. FRI layers of <= 2^10 values are committed by one workgroup in a
// single launch (k_fri_tail). Blake2s is VALU-bound here.
#include "vx_blake2s_x4.h"
#include "vx_channel_x4.h"
#include "vx_d2.h"
#include "vx_cairo.h"

namespace vx {

static constexpr uint32_t SUB_THREADS = 256;
static constexpr uint32_t BIG_TREE_LOG = 21;  // >= this: 16 leaves per lane, 4096-leaf tiles
static constexpr uint32_t SMALL_TREE_LOG = 16;  // <= this: 1 leaf per lane, 256-leaf tiles
static constexpr uint32_t MAX_TILE_LOG = 12;
static constexpr uint32_t RED_THREADS = 512;  // tree-top reduce: 1024 nodes, 10 levels per workgroup
static constexpr uint32_t RED_NLEV_MAX = 10;
static constexpr uint32_t RED_QUAD_NLEV = 8;  // 4 x 2^(8-1) lanes: the first level by quads
static constexpr uint32_t TAIL_THREADS = 1u << TAIL_MAX_LOG;  // one lane per leaf of the largest tail layer

struct G4 {
  uint32_t w[4][4];  // words of 4 consecutive leaves (W <= 4)
};
// Raw loads of one 4-leaf group (kept apart from the arithmetic so the next
// group's loads can be in flight while this group is hashed).
struct Raw {
  uint4 a[8];
  uint4 t;
  uint32_t h0, h1;
};

__device__ __forceinline__ uint4 ld4(const uint32_t* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ void st4(uint32_t* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  *reinterpret_cast<uint4*>(p) = make_uint4(a, b, c, d);
}
__device__ __forceinline__ qm31 ldq(const uint32_t* p) { return {{p[0], p[1], p[2], p[3]}}; }
__device__ __forceinline__ void u4w(const uint4& v, uint32_t* w) {
  w[0] = v.x;
  w[1] = v.y;
  w[2] = v.z;
  w[3] = v.w;
}

// Fibonacci constraint at position k of a window e[] (cuda/constraint.cu:22-30).
__device__ __forceinline__ uint32_t fib_c(const uint32_t* e, int k) {
  return m31_sub(e[k + 2], m31_add(e[k + 1], e[k]));
}

template <uint32_t KIND, uint32_t NC>
__device__ __forceinline__ void leaf_load(const LeafSrc& L, uint64_t i, Raw& r) {
  if constexpr (KIND == LEAF_COLS && NC == 1) {
    r.a[0] = ld4(L.col[0] + i);
  } else if constexpr (KIND == LEAF_COLS && NC == 4) {
#pragma unroll
    for (int c = 0; c < 4; c++) r.a[c] = ld4(L.col[c] + i);
  } else if constexpr (KIND == LEAF_QUOTIENT) {
    r.a[0] = ld4(L.col[0] + i);
    r.h0 = L.col[0][(i + 4) & L.src_mask];
    r.h1 = L.col[0][(i + 5) & L.src_mask];
  } else if constexpr (KIND == LEAF_FOLD_CIRCLE_Q) {
    const uint64_t b = 2 * i;
    r.a[0] = ld4(L.col[0] + b);
    r.a[1] = ld4(L.col[0] + b + 4);
    r.h0 = L.col[0][(b + 8) & L.src_mask];
    r.h1 = L.col[0][(b + 9) & L.src_mask];
    r.t = ld4(L.tw + i);
  } else if constexpr (KIND == LEAF_FOLD_LINE) {
    const uint64_t b = 2 * i;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      r.a[2 * c] = ld4(L.col[c] + b);
      r.a[2 * c + 1] = ld4(L.col[c] + b + 4);
    }
    r.t = ld4(L.tw + i);
  }
}

// Words of leaves [i, i+4) from their raw loads (fold producers also store
// the new layer).
template <uint32_t KIND, uint32_t NC>
__device__ __forceinline__ void leaf_produce(const LeafSrc& L, uint64_t i, const Raw& r, const qm31& aq,
                                             const qm31& af, const qm31_mat& am, G4& g, bool st = true) {
  if constexpr (KIND == LEAF_COLS && NC == 1) {
    g.w[0][0] = r.a[0].x;
    g.w[1][0] = r.a[0].y;
    g.w[2][0] = r.a[0].z;
    g.w[3][0] = r.a[0].w;
  } else if constexpr (KIND == LEAF_COLS && NC == 4) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      g.w[0][c] = r.a[c].x;
      g.w[1][c] = r.a[c].y;
      g.w[2][c] = r.a[c].z;
      g.w[3][c] = r.a[c].w;
    }
  } else if constexpr (KIND == LEAF_QUOTIENT) {
    uint32_t e[6];
    u4w(r.a[0], e);
    e[4] = r.h0;
    e[5] = r.h1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qm31 q = qm_scale(aq, fib_c(e, j));
#pragma unroll
      for (int c = 0; c < 4; c++) g.w[j][c] = q.v[c];
    }
  } else if constexpr (KIND == LEAF_FOLD_CIRCLE_Q) {
    // f_k = aq * c_k, so (f0 + f1) + af * (tw * (f0 - f1))
    //   = aq * (c0 + c1) + (af * aq) * (tw * (c0 - c1))   (exact field identity)
    uint32_t e[10], t[4];
    u4w(r.a[0], e);
    u4w(r.a[1], e + 4);
    e[8] = r.h0;
    e[9] = r.h1;
    u4w(r.t, t);
#pragma unroll
    for (int j = 0; j < 4; j++) {  // `af` holds af * aq here (see lane_subtree)
      const uint32_t c0 = fib_c(e, 2 * j), c1 = fib_c(e, 2 * j + 1);
      qm31 o = qm_add(qm_scale(aq, m31_add(c0, c1)), qm_scale(af, m31_mul(t[j], m31_sub(c0, c1))));
#pragma unroll
      for (int c = 0; c < 4; c++) g.w[j][c] = o.v[c];
    }
    if (st) {
#pragma unroll
      for (int c = 0; c < 4; c++) st4(L.out[c] + i, g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
    }
  } else if constexpr (KIND == LEAF_FOLD_LINE) {
    uint32_t in[4][8], t[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      u4w(r.a[2 * c], in[c]);
      u4w(r.a[2 * c + 1], in[c] + 4);
    }
    u4w(r.t, t);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qm31 f0 = {{in[0][2 * j], in[1][2 * j], in[2][2 * j], in[3][2 * j]}};
      qm31 f1 = {{in[0][2 * j + 1], in[1][2 * j + 1], in[2][2 * j + 1], in[3][2 * j + 1]}};
      qm31 o = fri_fold_pair_m(f0, f1, am, t[j]);
#pragma unroll
      for (int c = 0; c < 4; c++) g.w[j][c] = o.v[c];
    }
    if (st) {
#pragma unroll
      for (int c = 0; c < 4; c++) st4(L.out[c] + i, g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
    }
  }
}

// Synthetic AIR quotient of rows [i, i+4) (vx_d2.h):
//   Q(r) = sum_k gamma_k t_a[r] t_b[r+1] + delta_k t_e[r],
// gamma_k = beta_k c_k and delta_k = beta_k d_k (QM31, precomputed once per
// proof by d2_weights). Each QM31 component is a dot product accumulated in
// 64 bits without reduction: every product is below P^2, an accumulator
// folded to lo + 2 hi (below 3 * 2^32, congruent mod P since 2^32 = 2 mod P)
// takes four more products below 2^64, so it is folded after every pair of
// constraints and reduced once at the end. The constraint table and weights
// are wave-uniform (scalar loads); the rows are one 16-byte load per column
// use plus the wrapped row i + 4 of column b. Q is stored (SoA) for the fold.
__device__ __forceinline__ uint64_t fold64(uint64_t x) { return (x & 0xFFFFFFFFull) + ((x >> 32) << 1); }

__device__ __forceinline__ void d2_produce(const LeafSrc& L, uint64_t i, G4& g, bool store = true) {
  const uint32_t* ev = L.col[0];
  const uint64_t st = L.col_stride;
  const uint64_t i4 = (i + 4) & L.src_mask;
  const uint32_t K = L.n_cons;
  uint64_t acc[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[j][c] = 0;
#pragma unroll 2
  for (uint32_t k0 = 0; k0 < K; k0 += 2) {
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
      const uint32_t k = k0 + h;
      if (k < K) {
        const uint32_t* t = L.cons + 8 * k;
        const uint4 wg = ld_uniform(reinterpret_cast<const uint4*>(L.beta + 8 * k));
        const uint4 wd = ld_uniform(reinterpret_cast<const uint4*>(L.beta + 8 * k + 4));
        const uint32_t w[8] = {wg.x, wg.y, wg.z, wg.w, wd.x, wd.y, wd.z, wd.w};
        const uint32_t* cb = ev + ld_uniform(t + 3) * st;
        const uint4 va = ld4(ev + ld_uniform(t + 2) * st + i), vb = ld4(cb + i);
        const uint4 ve = ld4(ev + ld_uniform(t + 4) * st + i);
        const uint32_t b4 = cb[i4];
        const uint32_t A[4] = {va.x, va.y, va.z, va.w}, B[5] = {vb.x, vb.y, vb.z, vb.w, b4};
        const uint32_t E[4] = {ve.x, ve.y, ve.z, ve.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint64_t ab = m31_mul(A[j], B[j + 1]);
#pragma unroll
          for (int c = 0; c < 4; c++) acc[j][c] += (uint64_t)w[c] * ab + (uint64_t)w[4 + c] * E[j];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[j][c] = fold64(acc[j][c]);
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) g.w[j][c] = d2_mod_p(acc[j][c]);
  if (store) {
#pragma unroll
    for (int c = 0; c < 4; c++) st4(L.out[c] + i, g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
  }
}

// Lifted leaf (cuda/merkle_leaves_lifted.cu:26-98): Blake2s of the words of
// every column, in column order (ascending size), where a column of log
// size L < S contributes its row ((idx >> (S - L + 1)) << 1) | (idx & 1).
// 16 words per compression; the last block is the final one.
__device__ __forceinline__ uint64_t lifted_row(uint64_t idx, uint32_t S, uint32_t L) {
  return L == S ? idx : (((idx >> (S - L + 1)) << 1) | (idx & 1));
}
// Word at a 32-bit byte offset from a wave-uniform global base: one global
// load with a scalar base and a VGPR offset, no 64-bit address arithmetic.
__device__ __forceinline__ uint32_t ldg_at(const uint32_t* base, uint32_t byte_off) {
  const __attribute__((address_space(1))) char* b = (const __attribute__((address_space(1))) char*)base;
  return *(const __attribute__((address_space(1))) uint32_t*)(b + byte_off);
}
// The 16 message words of columns [base, base + 16) at leaf i's lifted rows
// (zero past the last column). Columns come sorted by size, so the row's
// byte offset changes only where the size does (9 times in a 96-column
// batch): the update stays a uniform branch, and a column otherwise costs
// one scalar table load and one global load. (As a select on every column
// the update cost 4 VALU instructions per column, about 9 % of the lifted
// tree's instructions; the empty asm statement keeps the compiler from
// turning the branch into selects.)
struct LiftCursor {
  uint32_t log, off;
};
__device__ __forceinline__ void lifted_block(const LiftCol* cols, uint32_t n, uint32_t S, uint32_t i, uint32_t base,
                                             LiftCursor& cur, uint32_t (&m)[16]) {
#pragma unroll
  for (uint32_t j = 0; j < 16; j++) {
    const uint32_t c = base + j;
    m[j] = 0;
    if (c < n) {
      const LiftCol col = ld_uniform(cols + c);
      if (col.log_n != cur.log) {
        asm volatile("" ::: "memory");
        cur.log = col.log_n;
        cur.off = col.log_n == S ? i << 2 : (((i >> (S - col.log_n + 1)) << 3) | ((i & 1u) << 2));
      }
      m[j] = ldg_at(col.p, cur.off);
    }
  }
}
__device__ __forceinline__ void lifted_leaf_cols(const LiftCol* cols, uint32_t n, uint32_t S, uint64_t idx,
                                                 uint32_t h[8]) {
  b2_init(h);
  if (S <= 30) {
    // rows and byte offsets fit 32 bits. (Loading the next block before
    // compressing this one takes the kernel from 70 to about 250 VGPRs.)
    const uint32_t i = (uint32_t)idx;
    LiftCursor cur{~0u, 0u};
    for (uint32_t base = 0; base < n; base += 16) {
      uint32_t m[16];
      const bool more = base + 16 < n;
      lifted_block(cols, n, S, i, base, cur, m);
      b2_compress(h, m, 4 * (more ? base + 16 : n), !more);
    }
    return;
  }
  uint32_t t = 0;
  for (uint32_t base = 0; base < n; base += 16) {
    uint32_t m[16];
#pragma unroll
    for (uint32_t j = 0; j < 16; j++) {
      const uint32_t c = base + j;
      uint32_t v = 0;
      if (c < n) {
        const LiftCol col = ld_uniform(cols + c);
        v = as_global(col.p)[lifted_row(idx, S, col.log_n)];
      }
      m[j] = v;
    }
    t += 4 * (n - base < 16 ? n - base : 16);
    b2_compress(h, m, t, base + 16 >= n);
  }
}
__device__ __forceinline__ void lifted_leaf(const LeafSrc& L, uint64_t idx, uint32_t h[8]) {
  lifted_leaf_cols(L.lift, L.lift_n, L.lift_log, idx, h);
}

// Random linear combination of the lifted rows of leaves [i, i+4):
// Q(r) = sum_c beta_c col_c[lifted_row(r)]. A column of the tree's size
// gives rows i..i+3 (one 16-byte load); a smaller one gives rows
// b, b+1, b, b+1 with b = (i >> shift) << 1 (one 8-byte load). The QM31
// components accumulate in 64 bits as in d2_produce (fold after every four
// columns); four columns' loads are in flight together. Q is stored (SoA)
// for the fold.
// Two words at a 32-bit byte offset from a wave-uniform global base.
__device__ __forceinline__ uint2 ldg2_at(const uint32_t* base, uint32_t byte_off) {
  const __attribute__((address_space(1))) char* b = (const __attribute__((address_space(1))) char*)base;
  const __attribute__((address_space(1))) uint32_t* q = (const __attribute__((address_space(1))) uint32_t*)(b + byte_off);
  return make_uint2(q[0], q[1]);
}
__device__ __forceinline__ void lifted_rlc_produce(const LeafSrc& L, uint64_t i, G4& g, bool store = true) {
  const uint32_t S = L.lift_log, n = L.lift_n;
  uint64_t acc[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[j][c] = 0;
  if (S <= 30) {
    // Rows and byte offsets fit 32 bits. Every column is two 8-byte loads
    // (rows i, i+1 and i+2, i+3 of a column of the tree's size; rows b,
    // b+1 twice of a smaller one) with the offsets picked by the uniform
    // size test, so the four columns' loads of a group carry no branch and
    // are all in flight before the first product.
    const uint32_t i32 = (uint32_t)i;
    for (uint32_t c0 = 0; c0 < n; c0 += 4) {
      uint2 lo[4], hi[4];
      uint4 bq[4];
#pragma unroll
      for (uint32_t h = 0; h < 4; h++) {
        const uint32_t c = c0 + h < n ? c0 + h : n - 1;  // past the end: a column whose products are dropped
        const LiftCol col = ld_uniform(L.lift + c);
        bq[h] = c0 + h < n ? ld_uniform(reinterpret_cast<const uint4*>(L.beta + 4 * c)) : make_uint4(0, 0, 0, 0);
        const bool full = col.log_n == S;
        const uint32_t b = (i32 >> (S - col.log_n + 1)) << 3;
        lo[h] = ldg2_at(col.p, full ? i32 << 2 : b);
        hi[h] = ldg2_at(col.p, full ? (i32 << 2) + 8 : b);
      }
#pragma unroll
      for (uint32_t h = 0; h < 4; h++) {
        const uint32_t be[4] = {bq[h].x, bq[h].y, bq[h].z, bq[h].w};
        const uint32_t v[4] = {lo[h].x, lo[h].y, hi[h].x, hi[h].y};
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int k = 0; k < 4; k++) acc[j][k] += (uint64_t)be[k] * v[j];
      }
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) acc[j][k] = fold64(acc[j][k]);
    }
  } else
  for (uint32_t c0 = 0; c0 < n; c0 += 4) {
#pragma unroll
    for (uint32_t h = 0; h < 4; h++) {
      const uint32_t c = c0 + h;
      if (c < n) {
        const LiftCol col = ld_uniform(L.lift + c);
        const uint4 bq = ld_uniform(reinterpret_cast<const uint4*>(L.beta + 4 * c));
        const uint32_t be[4] = {bq.x, bq.y, bq.z, bq.w};
        uint32_t v[4];
        if (col.log_n == S) {
          u4w(ld4_global(col.p + i), v);
        } else {
          const uint64_t b = (i >> (S - col.log_n + 1)) << 1;
          const uint2 x = ld2_global(col.p + b);
          v[0] = v[2] = x.x;
          v[1] = v[3] = x.y;
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int k = 0; k < 4; k++) acc[j][k] += (uint64_t)be[k] * v[j];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int k = 0; k < 4; k++) acc[j][k] = fold64(acc[j][k]);
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) g.w[j][c] = d2_mod_p(acc[j][c]);
  if (store) {
#pragma unroll
    for (int c = 0; c < 4; c++) st4(L.out[c] + i, g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
  }
}

// Rows [i, i+4) of the tile's quotient, written by the workgroup's first
// phase (k_subtree), read back from the output columns (L2-resident).
__device__ __forceinline__ void cairo_readback(const LeafSrc& L, uint64_t i, G4& g) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const uint4 v = ld4(L.out[c] + i);
    g.w[0][c] = v.x;
    g.w[1][c] = v.y;
    g.w[2][c] = v.z;
    g.w[3][c] = v.w;
  }
}

// One leaf per lane: the lane's own row, stored and hashed.
__device__ __forceinline__ void cairo_leaf(const LeafSrc& L, uint64_t idx, uint32_t out[8]) {
  const qm31 q = cairo_row(as_uniform(L.cairo), idx);
#pragma unroll
  for (int c = 0; c < 4; c++) L.out[c][idx] = q.v[c];
  b2_leaf4(q.v[0], q.v[1], q.v[2], q.v[3], out);
}

// Rows [i, i+4) of up to 16 columns, one 16-byte load per column, all
// issued before the first of the four leaves is hashed.
struct GCols {
  uint4 c[16];
};
__device__ __forceinline__ void cols_load(const LeafSrc& L, uint64_t i, GCols& gc) {
#pragma unroll
  for (uint32_t c = 0; c < 16; c++) gc.c[c] = c < L.n_cols ? ld4(L.col[c] + i) : make_uint4(0, 0, 0, 0);
}

template <uint32_t KIND, uint32_t NC>
__device__ __forceinline__ void leaf_hash(const LeafSrc& L, const G4& g, const GCols& gc, uint64_t i, int j,
                                          uint32_t h[8]) {
  if constexpr (KIND == LEAF_LIFTED) {
    lifted_leaf(L, i + j, h);
  } else if constexpr (KIND == LEAF_COLS && NC == 1) {
    b2_leaf1(g.w[j][0], h);
  } else if constexpr (KIND == LEAF_COLS && NC == 0) {
    // generic 1..16 columns: row i + j of each
    uint32_t w[16];
#pragma unroll
    for (uint32_t c = 0; c < 16; c++) w[c] = j == 0 ? gc.c[c].x : j == 1 ? gc.c[c].y : j == 2 ? gc.c[c].z : gc.c[c].w;
    b2_hash_words(w, L.n_cols, h);
  } else {
    b2_leaf4(g.w[j][0], g.w[j][1], g.w[j][2], g.w[j][3], h);
  }
}

__device__ __forceinline__ void cp8(uint32_t d[8], const uint32_t s[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) d[k] = s[k];
}

// Merkle node over the 2^LPT_LOG consecutive leaves starting at `first`.
template <uint32_t KIND, uint32_t NC, uint32_t LPT_LOG>
__device__ __forceinline__ void lane_subtree(const LeafSrc& L, uint64_t first, uint32_t out[8]) {
  static_assert(LPT_LOG >= 2 && LPT_LOG <= 4, "lane subtree covers 4..16 leaves");
  constexpr bool kGeneric = KIND == LEAF_COLS && NC == 0;
  constexpr bool kDirect = kGeneric || KIND == LEAF_D2_QUOTIENT || KIND == LEAF_LIFTED ||
                           KIND == LEAF_LIFTED_RLC || KIND == LEAF_CAIRO_QUOTIENT;  // no raw-load prefetch
  constexpr uint32_t G = 1u << (LPT_LOG - 2);
  qm31 aq{}, af{};
  qm31_mat am{};
  if constexpr (KIND == LEAF_QUOTIENT || KIND == LEAF_FOLD_CIRCLE_Q) aq = ldq(L.alpha_q);
  if constexpr (KIND == LEAF_FOLD_LINE) am = qm_left_matrix(ld_uniform(reinterpret_cast<const qm31*>(L.alpha_f)));
  if constexpr (KIND == LEAF_FOLD_CIRCLE_Q) af = qm_mul(ldq(L.alpha_f), aq);
  Raw cur, nxt;
  if constexpr (!kDirect) leaf_load<KIND, NC>(L, first, cur);
  uint32_t s2[8], s3[8];
  // Not unrolled: a group's body is seven compressions (about 7,000
  // instructions), so two or four copies would exceed what the instruction
  // cache holds well; the branches on gi are uniform.
#pragma unroll 1
  for (uint32_t gi = 0; gi < G; gi++) {
    const uint64_t i = first + 4 * gi;
    G4 g;
    GCols gc;
    if constexpr (!kDirect) {
      if (gi + 1 < G) leaf_load<KIND, NC>(L, i + 4, nxt);
      leaf_produce<KIND, NC>(L, i, cur, aq, af, am, g);
    }
    if constexpr (kGeneric) cols_load(L, i, gc);
    if constexpr (KIND == LEAF_D2_QUOTIENT) d2_produce(L, i, g);
    if constexpr (KIND == LEAF_LIFTED_RLC) lifted_rlc_produce(L, i, g);
    if constexpr (KIND == LEAF_CAIRO_QUOTIENT) cairo_readback(L, i, g);
    uint32_t a[8], b[8], n1a[8], n1b[8], n2[8];
    leaf_hash<KIND, NC>(L, g, gc, i, 0, a);
    leaf_hash<KIND, NC>(L, g, gc, i, 1, b);
    b2_hash_node(a, b, n1a);
    leaf_hash<KIND, NC>(L, g, gc, i, 2, a);
    leaf_hash<KIND, NC>(L, g, gc, i, 3, b);
    b2_hash_node(a, b, n1b);
    b2_hash_node(n1a, n1b, n2);
    if constexpr (LPT_LOG == 2) {
      cp8(out, n2);
    } else {
      if ((gi & 1) == 0) {
        cp8(s2, n2);
      } else {
        uint32_t n3[8];
        b2_hash_node(s2, n2, n3);
        if constexpr (LPT_LOG == 3) {
          cp8(out, n3);
        } else {
          if ((gi & 2) == 0) cp8(s3, n3);
          else b2_hash_node(s3, n3, out);
        }
      }
    }
    if constexpr (!kDirect) {
      if (gi + 1 < G) cur = nxt;
    }
  }
}

// One leaf per lane (small, latency-bound trees): the producers work on
// groups of four leaves, so each lane computes the words of its group (the
// group's four lanes load the same addresses) and hashes leaf idx & 3; only
// the group's first lane stores produced values (fold layer, quotient).
template <uint32_t KIND, uint32_t NC>
__device__ __forceinline__ void lane_leaf(const LeafSrc& L, uint64_t idx, uint32_t out[8]) {
  constexpr bool kGeneric = KIND == LEAF_COLS && NC == 0;
  if constexpr (KIND == LEAF_LIFTED) {
    lifted_leaf(L, idx, out);
  } else if constexpr (KIND == LEAF_CAIRO_QUOTIENT) {
    cairo_leaf(L, idx, out);
  } else {
    const uint64_t i = idx & ~3ull;
    const uint32_t j = (uint32_t)idx & 3u;
    const bool first = j == 0;
    auto pick = [j](uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
      return j == 0 ? a0 : (j == 1 ? a1 : (j == 2 ? a2 : a3));
    };
    G4 g;
    if constexpr (kGeneric) {
      GCols gc;
      cols_load(L, i, gc);
      uint32_t w[16];
#pragma unroll
      for (uint32_t c = 0; c < 16; c++) w[c] = pick(gc.c[c].x, gc.c[c].y, gc.c[c].z, gc.c[c].w);
      b2_hash_words(w, L.n_cols, out);
      return;
    } else if constexpr (KIND == LEAF_D2_QUOTIENT) {
      d2_produce(L, i, g, first);
    } else if constexpr (KIND == LEAF_LIFTED_RLC) {
      lifted_rlc_produce(L, i, g, first);
    } else {
      qm31 aq{}, af{};
      qm31_mat am{};
      if constexpr (KIND == LEAF_QUOTIENT || KIND == LEAF_FOLD_CIRCLE_Q) aq = ldq(L.alpha_q);
      if constexpr (KIND == LEAF_FOLD_LINE)
        am = qm_left_matrix(ld_uniform(reinterpret_cast<const qm31*>(L.alpha_f)));
      if constexpr (KIND == LEAF_FOLD_CIRCLE_Q) af = qm_mul(ldq(L.alpha_f), aq);
      Raw r;
      leaf_load<KIND, NC>(L, i, r);
      leaf_produce<KIND, NC>(L, i, r, aq, af, am, g, first);
    }
    if constexpr (KIND == LEAF_COLS && NC == 1) {
      b2_leaf1(pick(g.w[0][0], g.w[1][0], g.w[2][0], g.w[3][0]), out);
    } else {
      uint32_t w[4];
#pragma unroll
      for (int c = 0; c < 4; c++) w[c] = pick(g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
      b2_leaf4(w[0], w[1], w[2], w[3], out);
    }
  }
}

__device__ __forceinline__ void store_node(uint32_t* p, const uint32_t n[8]) {
  st4(p, n[0], n[1], n[2], n[3]);
  st4(p + 4, n[4], n[5], n[6], n[7]);
}
__device__ __forceinline__ void lds_put(uint4* s, uint32_t idx, const uint32_t n[8]) {
  s[2 * idx] = make_uint4(n[0], n[1], n[2], n[3]);
  s[2 * idx + 1] = make_uint4(n[4], n[5], n[6], n[7]);
}
__device__ __forceinline__ void lds_get(const uint4* s, uint32_t idx, uint32_t n[8]) {
  uint4 a = s[2 * idx], b = s[2 * idx + 1];
  n[0] = a.x; n[1] = a.y; n[2] = a.z; n[3] = a.w;
  n[4] = b.x; n[5] = b.y; n[6] = b.z; n[7] = b.w;
}

// Transcript step fused into the kernel that produced a root
// (src/prover.rs:321-334 and the mix/draw pairs that follow each commit),
// run by the quad of lanes 0..3, which all hold the root: the channel's
// compressions are quad hashes (about half the dependent chain of one lane).
// Lane 0 writes the results.
__device__ __forceinline__ void finish_root_q0(const RootHook& h, const uint32_t node[8], uint32_t q) {
  if (h.root_out && q == 0)
    for (int k = 0; k < 8; k++) h.root_out[k] = node[k];
  if (h.ch) {
    Channel c;
    if (h.init) ch_init(c);
    else c = *h.ch;
    ch_mix_digest_q0(c, node, q);
    if (h.alpha_out) {
      const qm31 a = ch_draw_felt_q0(c, q);
      if (q == 0)
        for (int k = 0; k < 4; k++) h.alpha_out[k] = a.v[k];
    }
    if (q == 0) *h.ch = c;
  }
}

// Merges `nthr` nodes held one per lane in `buf[0]` up to one node, storing
// each produced level (lvl0+1 ...) at global index base_idx(level) + lane.
// Returns the top node in lane 0's `node`.
// With QUAD, levels with at most CAP / 4 nodes hash each node with the four
// lanes of a quad (b2_node_x4): those levels are a dependent chain with most
// lanes idle, so halving the instructions per compression halves their
// latency. (The throughput instances do without: the quad path raises their
// VGPR count.) Lanes 0..3 return the top node.
template <uint32_t CAP, bool QUAD>
__device__ __forceinline__ void lds_merge(uint4 (*buf)[2 * CAP], uint32_t nthr, uint32_t lvl0,
                                          const TreeLevels& T, uint64_t group, uint32_t group_log,
                                          uint32_t store_lo, uint32_t node[8]) {
  const uint32_t tid = threadIdx.x;
  uint32_t cur = 0;
  uint32_t cnt = nthr;
  for (uint32_t lvl = lvl0 + 1; cnt > 1; lvl++) {
    cnt >>= 1;
    if (QUAD && 4 * cnt <= CAP) {
      const uint32_t nid = tid >> 2, q = tid & 3;
      if (nid < cnt) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(buf[cur]) + 16 * nid;
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = src[k];
        uint32_t lo, hi;
        b2_node_x4(m, q, lo, hi);
        uint32_t* dst = reinterpret_cast<uint32_t*>(buf[cur ^ 1]) + 8 * nid;
        dst[q] = lo;
        dst[4 + q] = hi;
        if (T.base && lvl >= store_lo) {
          uint32_t* g = T.base + T.off[lvl] + 8 * ((group << (group_log - lvl)) + nid);
          g[q] = lo;
          g[4 + q] = hi;
        }
      }
    } else if (tid < cnt) {
      uint32_t l[8], r[8];
      lds_get(buf[cur], 2 * tid, l);
      lds_get(buf[cur], 2 * tid + 1, r);
      b2_hash_node(l, r, node);
      lds_put(buf[cur ^ 1], tid, node);
      if (T.base && lvl >= store_lo)
        store_node(T.base + T.off[lvl] + 8 * ((group << (group_log - lvl)) + tid), node);
    }
    cur ^= 1;
    __syncthreads();
  }
  if (tid < 4) lds_get(buf[cur], 0, node);
}

// ROOT instances (the whole tree in one workgroup) carry the transcript step;
// the throughput instances do not, which keeps their VGPR count (and so their
// occupancy) at what the hashing itself needs.
template <uint32_t KIND, uint32_t NC, uint32_t LPT_LOG, bool ROOT>
__global__ __launch_bounds__(SUB_THREADS) void k_subtree(LeafSrc L, TreeLevels T, uint32_t tile_log,
                                                         uint32_t store_lo, RootHook hook) {
  __shared__ uint4 buf[2][2 * SUB_THREADS];
  const uint32_t nthr = 1u << (tile_log - LPT_LOG);
  const uint32_t tid = threadIdx.x;
  const uint64_t tile = blockIdx.x;
  uint32_t node[8];
  if constexpr (KIND == LEAF_CAIRO_QUOTIENT && LPT_LOG >= 2) {
    // First phase: the tile's quotient rows one per lane, so consecutive
    // lanes read consecutive rows of every trace column (a lane computing
    // its own 2^LPT_LOG rows would stride them); stored for the FRI, then
    // read back by the hashing lanes after the barrier.
    uniform_t<CairoArgs>& a = as_uniform(L.cairo);
    const uint64_t base = tile << tile_log;
#pragma unroll 1
    for (uint32_t r = tid; r < (1u << tile_log); r += SUB_THREADS) {
      const qm31 q = cairo_row(a, base + r);
#pragma unroll
      for (int c = 0; c < 4; c++) L.out[c][base + r] = q.v[c];
    }
    __syncthreads();
  }
  if (tid < nthr) {
    const uint64_t first = (tile << tile_log) + ((uint64_t)tid << LPT_LOG);
    if constexpr (LPT_LOG == 0) lane_leaf<KIND, NC>(L, first, node);
    else lane_subtree<KIND, NC, LPT_LOG>(L, first, node);
    lds_put(buf[0], tid, node);
    if (T.base && LPT_LOG >= store_lo)
      store_node(T.base + T.off[LPT_LOG] + 8 * ((tile << (tile_log - LPT_LOG)) + tid), node);
  }
  __syncthreads();
  lds_merge<SUB_THREADS, ROOT || LPT_LOG <= 2>(buf, nthr, LPT_LOG, T, tile, tile_log, store_lo, node);
  if constexpr (ROOT) {
    if (tid < 4) finish_root_q0(hook, node, tid);
  }
}

// Upper tree levels: each workgroup hashes 2^nlev level-`from` nodes down to
// one node, storing levels from+1 .. from+nlev.
template <bool ROOT>
__global__ __launch_bounds__(RED_THREADS) void k_tree_reduce(TreeLevels T, uint32_t from, uint32_t nlev,
                                                             RootHook hook) {
  __shared__ uint4 buf[2][2 * RED_THREADS];
  const uint32_t tid = threadIdx.x;
  const uint32_t nthr = 1u << (nlev - 1);
  const uint64_t wg = blockIdx.x;
  uint32_t node[8];
  if (4 * nthr <= RED_THREADS) {  // few nodes: the first level by quads too
    const uint32_t nid = tid >> 2, q = tid & 3;
    if (nid < nthr) {
      const uint32_t* in = T.base + T.off[from] + 8 * ((wg << nlev) + 2 * (uint64_t)nid);
      uint32_t m[16];
#pragma unroll
      for (int k = 0; k < 16; k++) m[k] = in[k];
      uint32_t lo, hi;
      b2_node_x4(m, q, lo, hi);
      uint32_t* dst = reinterpret_cast<uint32_t*>(buf[0]) + 8 * nid;
      dst[q] = lo;
      dst[4 + q] = hi;
      uint32_t* g = T.base + T.off[from + 1] + 8 * ((wg << (nlev - 1)) + nid);
      g[q] = lo;
      g[4 + q] = hi;
    }
  } else if (tid < nthr) {
    const uint32_t* in = T.base + T.off[from] + 8 * ((wg << nlev) + 2 * (uint64_t)tid);
    uint32_t l[8], r[8];
    for (int k = 0; k < 8; k++) {
      l[k] = in[k];
      r[k] = in[8 + k];
    }
    b2_hash_node(l, r, node);
    lds_put(buf[0], tid, node);
    store_node(T.base + T.off[from + 1] + 8 * ((wg << (nlev - 1)) + tid), node);
  }
  __syncthreads();
  lds_merge<RED_THREADS, true>(buf, nthr, from + 1, T, wg, from + nlev, 0, node);
  if constexpr (ROOT) {
    if (tid < 4) finish_root_q0(hook, node, tid);
  }
}

// Leaves per lane (log) for a tile of 2^tile_log leaves: 256 lanes when the
// tile allows it, between 4 and 16 leaves each.
static uint32_t lane_log_for(uint32_t tile_log) {
  uint32_t l = tile_log > 8 ? tile_log - 8 : 0;
  return l < 2 ? 2 : (l > 4 ? 4 : l);
}

template <uint32_t KIND, uint32_t NC, uint32_t LPT>
static void launch_subtree_lpt(const LeafSrc& L, const TreeLevels& T, uint32_t tile_log, uint32_t store_lo,
                               const RootHook& hook, hipStream_t s) {
  const uint64_t grid = 1ull << (T.log_n - tile_log);
  if (tile_log == T.log_n)
    hipLaunchKernelGGL((k_subtree<KIND, NC, LPT, true>), dim3((uint32_t)grid), dim3(SUB_THREADS), 0, s, L, T,
                       tile_log, store_lo, hook);
  else
    hipLaunchKernelGGL((k_subtree<KIND, NC, LPT, false>), dim3((uint32_t)grid), dim3(SUB_THREADS), 0, s, L, T,
                       tile_log, store_lo, hook);
}

template <uint32_t KIND, uint32_t NC>
static void launch_subtree(const LeafSrc& L, const TreeLevels& T, uint32_t tile_log, uint32_t lpt,
                           uint32_t store_lo, const RootHook& hook, hipStream_t s) {
  require(tile_log >= lpt && (tile_log - lpt) <= 8 && (lpt == 0 || (lpt >= 2 && lpt <= 4)),
          "merkle: tile shape out of range");
  if (lpt == 4) launch_subtree_lpt<KIND, NC, 4>(L, T, tile_log, store_lo, hook, s);
  else if (lpt == 3) launch_subtree_lpt<KIND, NC, 3>(L, T, tile_log, store_lo, hook, s);
  else if (lpt == 2) launch_subtree_lpt<KIND, NC, 2>(L, T, tile_log, store_lo, hook, s);
  else launch_subtree_lpt<KIND, NC, 0>(L, T, tile_log, store_lo, hook, s);
  VX_HIP(hipGetLastError());
}

static void check_leaf(const LeafSrc& L, uint32_t log_n) {
  require(log_n >= 2 && log_n <= 31, "merkle: log_n out of range");
  switch (L.kind) {
    case LEAF_COLS:
      require(L.n_cols >= 1 && L.n_cols <= 16, "merkle: n_cols must be 1..16");
      for (uint32_t c = 0; c < L.n_cols; c++)
        require(L.col[c] != nullptr && (reinterpret_cast<uintptr_t>(L.col[c]) & 15u) == 0,
                "merkle: columns must be non-null and 16-byte aligned");
      break;
    case LEAF_QUOTIENT:
      require(L.col[0] && L.alpha_q && L.src_mask == (1ull << log_n) - 1, "merkle: bad quotient source");
      break;
    case LEAF_FOLD_CIRCLE_Q:
      require(L.col[0] && L.alpha_q && L.alpha_f && L.tw && L.src_mask == (2ull << log_n) - 1,
              "merkle: bad circle-fold source");
      for (int c = 0; c < 4; c++) require(L.out[c] != nullptr, "merkle: null fold output");
      break;
    case LEAF_FOLD_LINE:
      require(L.alpha_f && L.tw, "merkle: bad line-fold source");
      for (int c = 0; c < 4; c++) require(L.col[c] && L.out[c], "merkle: null fold column");
      break;
    case LEAF_LIFTED:
      require(L.lift && L.lift_n >= 1 && L.lift_log == log_n, "merkle: bad lifted source");
      break;
    case LEAF_LIFTED_RLC:
      require(L.lift && L.lift_n >= 1 && L.lift_log == log_n && L.beta, "merkle: bad lifted combination source");
      for (int c = 0; c < 4; c++) require(L.out[c] != nullptr, "merkle: null quotient output");
      break;
    case LEAF_CAIRO_QUOTIENT:
      require(L.cairo != nullptr, "merkle: bad Cairo quotient source");
      for (int c = 0; c < 4; c++)
        require(L.out[c] != nullptr && (reinterpret_cast<uintptr_t>(L.out[c]) & 15u) == 0,
                "merkle: Cairo quotient outputs must be non-null and 16-byte aligned");
      break;
    case LEAF_D2_QUOTIENT:
      require(L.col[0] && L.cons && L.beta && L.n_cols >= 1 && L.n_cons >= 1 &&
                  L.src_mask == (1ull << log_n) - 1 && L.col_stride >= (1ull << log_n),
              "merkle: bad multi-column quotient source");
      for (int c = 0; c < 4; c++) require(L.out[c] != nullptr, "merkle: null quotient output");
      break;
    default:
      require(false, "merkle: unknown leaf kind");
  }
}

static void dispatch_subtree(const LeafSrc& L, const TreeLevels& T, uint32_t tile_log, uint32_t lpt,
                             uint32_t store_lo, const RootHook& hook, hipStream_t s) {
  switch (L.kind) {
    case LEAF_COLS:
      if (L.n_cols == 1) launch_subtree<LEAF_COLS, 1>(L, T, tile_log, lpt, store_lo, hook, s);
      else if (L.n_cols == 4) launch_subtree<LEAF_COLS, 4>(L, T, tile_log, lpt, store_lo, hook, s);
      else launch_subtree<LEAF_COLS, 0>(L, T, tile_log, lpt, store_lo, hook, s);
      break;
    case LEAF_QUOTIENT: launch_subtree<LEAF_QUOTIENT, 4>(L, T, tile_log, lpt, store_lo, hook, s); break;
    case LEAF_FOLD_CIRCLE_Q: launch_subtree<LEAF_FOLD_CIRCLE_Q, 4>(L, T, tile_log, lpt, store_lo, hook, s); break;
    case LEAF_FOLD_LINE: launch_subtree<LEAF_FOLD_LINE, 4>(L, T, tile_log, lpt, store_lo, hook, s); break;
    case LEAF_D2_QUOTIENT: launch_subtree<LEAF_D2_QUOTIENT, 4>(L, T, tile_log, lpt, store_lo, hook, s); break;
    case LEAF_LIFTED: launch_subtree<LEAF_LIFTED, 0>(L, T, tile_log, lpt, store_lo, hook, s); break;
    case LEAF_LIFTED_RLC: launch_subtree<LEAF_LIFTED_RLC, 4>(L, T, tile_log, lpt, store_lo, hook, s); break;
    case LEAF_CAIRO_QUOTIENT: launch_subtree<LEAF_CAIRO_QUOTIENT, 4>(L, T, tile_log, lpt, store_lo, hook, s); break;
  }
}

// Built-in tile (leaves per subtree workgroup, log) and leaves per lane
// (log) of a tree of 2^log_n leaves of the given kind.
TreeShape merkle_default_shape(const LeafSrc& L, uint32_t log_n) {
  // Leaves read from several columns (and the FRI folds, which read two
  // values of four columns per leaf) keep 4 leaves per lane at every size:
  // lane t then reads rows 4t..4t+3, so each 16-byte load of a wavefront is
  // one contiguous run. With 16 leaves per lane a wavefront's load would
  // touch 64 lines at a 256-byte (folds: 512-byte) stride and the rest of
  // each line is used only after the lane's next hash groups, by which time
  // L2 has evicted it (measured: line folds fetched each byte about 4 times).
  const bool multi = L.kind == LEAF_D2_QUOTIENT || L.kind == LEAF_LIFTED_RLC || L.kind == LEAF_LIFTED ||
                     L.kind == LEAF_CAIRO_QUOTIENT ||
                     (L.kind == LEAF_COLS && L.n_cols > 1) || L.kind == LEAF_FOLD_LINE ||
                     L.kind == LEAF_FOLD_CIRCLE_Q;
  // Small trees are latency-bound: one leaf per lane and 256-leaf tiles
  // make a workgroup's dependent chain one leaf hash plus the tile's levels
  // (quad hashes from 64 nodes down) instead of 7 hashes plus 8 levels.
  if (log_n <= SMALL_TREE_LOG) return {log_n < 8 ? log_n : 8, 0};
  const uint32_t tile_log = log_n >= BIG_TREE_LOG && !multi ? MAX_TILE_LOG : 10;
  return {tile_log, lane_log_for(tile_log)};
}

std::vector<TreeShape> merkle_shapes(uint32_t log_n, uint32_t store_lo) {
  std::vector<TreeShape> out;
  const TreeShape all[4] = {{log_n < 8 ? log_n : 8, 0}, {10, 2}, {11, 3}, {12, 4}};
  for (const TreeShape& t : all)
    if (t.tile_log <= log_n && t.lpt <= store_lo && store_lo <= t.tile_log) out.push_back(t);
  return out;
}

uint32_t merkle_tile_log(const LeafSrc& L, uint32_t log_n) { return merkle_default_shape(L, log_n).tile_log; }

void merkle_commit(const LeafSrc& L, const TreeLevels& T, uint32_t store_lo, const RootHook& hook, hipStream_t s,
                   hipEvent_t sub_begin, hipEvent_t sub_end) {
  merkle_commit_shaped(L, T, store_lo, hook, s, tuned_shape(L, T, store_lo, s), sub_begin, sub_end);
}

void merkle_commit_shaped(const LeafSrc& L, const TreeLevels& T, uint32_t store_lo, const RootHook& hook,
                          hipStream_t s, const TreeShape& shape, hipEvent_t sub_begin, hipEvent_t sub_end) {
  check_leaf(L, T.log_n);
  require(T.base != nullptr, "merkle_commit: null tree buffer");
  const uint32_t tile_log = shape.tile_log, lpt = shape.lpt;
  require(tile_log <= T.log_n && tile_log <= MAX_TILE_LOG && tile_log >= lpt && tile_log - lpt <= 8,
          "merkle_commit: invalid tree shape");
  require(store_lo >= T.lo && store_lo >= lpt && store_lo <= tile_log,
          "merkle_commit: stored levels must start at or above the lane level");
  if (sub_begin) VX_HIP(hipEventRecord(sub_begin, s));
  dispatch_subtree(L, T, tile_log, lpt, store_lo, hook, s);
  if (sub_end) VX_HIP(hipEventRecord(sub_end, s));
  // Levels above the tiles: up to RED_NLEV_MAX in one workgroup. More are
  // split evenly over launches of at most 8 levels, so that every workgroup
  // hashes its first level by quads (at most 128 nodes) and the lower
  // launches spread over many CUs: the levels are a dependent chain, and a
  // level of 512 nodes hashed one per lane on one CU takes about four quad
  // levels' time.
  const uint32_t above = T.log_n - tile_log;
  const uint32_t launches = above <= RED_NLEV_MAX ? 1 : (above + RED_QUAD_NLEV - 1) / RED_QUAD_NLEV;
  uint32_t from = tile_log;
  for (uint32_t k = 0; from < T.log_n; k++) {
    const uint32_t left = launches - k;  // levels of this launch: an even share of what is left
    const uint32_t nlev = (T.log_n - from + left - 1) / left;
    uint64_t grid = 1ull << (T.log_n - from - nlev);
    if (from + nlev == T.log_n)
      hipLaunchKernelGGL(k_tree_reduce<true>, dim3((uint32_t)grid), dim3(RED_THREADS), 0, s, T, from, nlev, hook);
    else
      hipLaunchKernelGGL(k_tree_reduce<false>, dim3((uint32_t)grid), dim3(RED_THREADS), 0, s, T, from, nlev, hook);
    VX_HIP(hipGetLastError());
    from += nlev;
  }
}

void merkle_tiled(const LeafSrc& L, uint32_t log_n, uint32_t tile_log, uint32_t* subtree_roots, hipStream_t s) {
  check_leaf(L, log_n);
  require(tile_log >= 4 && tile_log <= MAX_TILE_LOG && tile_log <= log_n, "merkle_tiled: tile size out of range");
  TreeLevels T{};
  T.base = subtree_roots;
  T.log_n = log_n;
  T.lo = tile_log;
  T.off[tile_log] = 0;
  dispatch_subtree(L, T, tile_log, lane_log_for(tile_log), tile_log, RootHook{}, s);
}

// ------------------------------------------------------------ FRI tail
// FRI line-fold layers of at most 2^TAIL_MAX_LOG values, committed one after
// another by a single 1024-lane workgroup: fold (previous layer from LDS),
// one leaf hash per lane, tree levels in LDS (quad hashes once at most 256
// parents are left; levels >= lo stored for the decommitment), root into the
// proof, then the quad of lanes 0..3 mixes the root and draws the next alpha
// (src/prover.rs:527-544). A layer of 2^L values is a dependent chain of
// two one-lane hashes, L - 1 quad hashes and the two quad channel hashes.
__global__ __launch_bounds__(TAIL_THREADS) void k_fri_tail(FriTailArgs a) {
  __shared__ uint32_t vals[2][4][1u << TAIL_MAX_LOG];
  __shared__ uint4 hashes[2u << TAIL_MAX_LOG];
  __shared__ uint32_t s_alpha[4];
  const uint32_t tid = threadIdx.x;
  uint32_t cur = 0;
  for (uint32_t li = 0; li < a.n_layers; li++) {
    const FriTailLayer& T = a.layer[li];
    const uint32_t L = T.log_n, n = 1u << L;
    if (tid == 0)
      for (int k = 0; k < 4; k++) s_alpha[k] = T.alpha[k];
    __syncthreads();
    const qm31_mat am = qm_left_matrix(qm31{{s_alpha[0], s_alpha[1], s_alpha[2], s_alpha[3]}});
    if (tid < n) {
      const uint32_t i = tid;
      qm31 f0, f1;
      if (li == 0) {
        f0 = {{a.in[0][2 * i], a.in[1][2 * i], a.in[2][2 * i], a.in[3][2 * i]}};
        f1 = {{a.in[0][2 * i + 1], a.in[1][2 * i + 1], a.in[2][2 * i + 1], a.in[3][2 * i + 1]}};
      } else {
        const uint32_t p = cur ^ 1;
        f0 = {{vals[p][0][2 * i], vals[p][1][2 * i], vals[p][2][2 * i], vals[p][3][2 * i]}};
        f1 = {{vals[p][0][2 * i + 1], vals[p][1][2 * i + 1], vals[p][2][2 * i + 1], vals[p][3][2 * i + 1]}};
      }
      const qm31 r = fri_fold_pair_m(f0, f1, am, T.tw[i]);
      for (int c = 0; c < 4; c++) {
        vals[cur][c][i] = r.v[c];
        T.out[c][i] = r.v[c];
      }
      uint32_t h[8];
      b2_leaf4(r.v[0], r.v[1], r.v[2], r.v[3], h);
      lds_put(hashes, i, h);
    }
    __syncthreads();
    for (uint32_t lvl = 1; lvl <= L; lvl++) {
      const uint32_t cnt = n >> lvl;
      if (4 * cnt <= TAIL_THREADS) {
        const uint32_t nid = tid >> 2, q = tid & 3;
        uint32_t lo = 0, hi = 0;
        if (nid < cnt) {
          const uint32_t* src = reinterpret_cast<const uint32_t*>(hashes) + 16 * nid;
          uint32_t m[16];
#pragma unroll
          for (int k = 0; k < 16; k++) m[k] = src[k];
          b2_node_x4(m, q, lo, hi);
        }
        __syncthreads();  // every child read before any parent overwrites it
        if (nid < cnt) {
          uint32_t* dst = reinterpret_cast<uint32_t*>(hashes) + 8 * nid;
          dst[q] = lo;
          dst[4 + q] = hi;
          if (lvl >= T.lo) {
            uint32_t* g = T.tree + T.off[lvl] + 8 * nid;
            g[q] = lo;
            g[4 + q] = hi;
          }
        }
      } else {
        uint32_t p[8], l[8], r[8];
        if (tid < cnt) {
          lds_get(hashes, 2 * tid, l);
          lds_get(hashes, 2 * tid + 1, r);
          b2_hash_node(l, r, p);
        }
        __syncthreads();
        if (tid < cnt) {
          lds_put(hashes, tid, p);
          if (lvl >= T.lo) store_node(T.tree + T.off[lvl] + 8 * tid, p);
        }
      }
      __syncthreads();
    }
    if (tid < 4) {
      uint32_t root[8];
      lds_get(hashes, 0, root);
      if (tid == 0)
        for (int k = 0; k < 8; k++) T.root_out[k] = root[k];
      Channel c = *a.ch;
      ch_mix_digest_q0(c, root, tid);
      if (T.next_alpha) {
        const qm31 nx = ch_draw_felt_q0(c, tid);
        if (tid == 0)
          for (int k = 0; k < 4; k++) T.next_alpha[k] = nx.v[k];
      }
      if (tid == 0) *a.ch = c;
    }
    __syncthreads();
    cur ^= 1;
  }
}

void fri_tail(const FriTailArgs& a, hipStream_t s) {
  require(a.n_layers >= 1 && a.n_layers <= TAIL_MAX_LAYERS && a.ch, "fri_tail: bad layer count");
  uint32_t prev = 0;
  for (uint32_t li = 0; li < a.n_layers; li++) {
    const FriTailLayer& T = a.layer[li];
    require(T.log_n >= 1 && T.log_n <= TAIL_MAX_LOG && T.lo <= T.log_n && T.tw && T.tree && T.root_out && T.alpha,
            "fri_tail: bad layer");
    require(li == 0 || T.log_n + 1 == prev, "fri_tail: layers must halve");
    for (int c = 0; c < 4; c++) require(T.out[c] != nullptr && (li > 0 || a.in[c] != nullptr), "fri_tail: null column");
    prev = T.log_n;
  }
  hipLaunchKernelGGL(k_fri_tail, dim3(1), dim3(TAIL_THREADS), 0, s, a);
  VX_HIP(hipGetLastError());
}

// ------------------------------------------------------- reference-API kernels
__global__ void k_leaf_hashes(LeafSrc L, uint64_t n, uint32_t* hashes) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t w[16];
  uint32_t nc = L.n_cols > 16 ? 16 : L.n_cols;
  for (uint32_t c = 0; c < 16; c++) w[c] = c < nc ? L.col[c][i] : 0u;
  uint32_t h[8];
  b2_hash_words(w, nc, h);
  store_node(hashes + 8 * i, h);
}
void merkle_leaf_hashes(const LeafSrc& L, uint64_t n, uint32_t* hashes, hipStream_t s) {
  require(L.kind == LEAF_COLS && L.n_cols >= 1 && L.n_cols <= 16, "merkle_hash_leaves: 1..16 columns");
  if (n == 0) return;
  hipLaunchKernelGGL(k_leaf_hashes, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, L, n, hashes);
  VX_HIP(hipGetLastError());
}

__global__ void k_hash_nodes(const uint32_t* ch, uint32_t* par, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t l[8], r[8], h[8];
  for (int k = 0; k < 8; k++) {
    l[k] = ch[16 * i + k];
    r[k] = ch[16 * i + 8 + k];
  }
  b2_hash_node(l, r, h);
  store_node(par + 8 * i, h);
}
void merkle_hash_nodes(const uint32_t* children, uint32_t* parents, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_hash_nodes, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, children, parents, n);
  VX_HIP(hipGetLastError());
}

// <= 1024 nodes (power of two) to a root in one workgroup
// (cuda/blake2s.cu:244-297 semantics).
static constexpr uint32_t RTR_THREADS = 512;
__global__ __launch_bounds__(RTR_THREADS) void k_reduce_to_root(const uint32_t* nodes, uint32_t* root,
                                                                uint32_t n_nodes) {
  __shared__ uint4 buf[2][2 * RTR_THREADS];
  const uint32_t tid = threadIdx.x;
  uint32_t node[8];
  if (n_nodes == 1) {
    if (tid < 8) root[tid] = nodes[tid];
    return;
  }
  const uint32_t nthr = n_nodes / 2;
  if (tid < nthr) {
    uint32_t l[8], r[8];
    for (int k = 0; k < 8; k++) {
      l[k] = nodes[16 * tid + k];
      r[k] = nodes[16 * tid + 8 + k];
    }
    b2_hash_node(l, r, node);
    lds_put(buf[0], tid, node);
  }
  __syncthreads();
  TreeLevels none{};
  lds_merge<RTR_THREADS, true>(buf, nthr, 0, none, 0, 0, 0, node);
  if (tid == 0)
    for (int k = 0; k < 8; k++) root[k] = node[k];
}
void merkle_reduce_to_root(const uint32_t* nodes, uint32_t* root, uint32_t n_nodes, hipStream_t s) {
  require(n_nodes >= 1 && n_nodes <= 2 * RTR_THREADS && (n_nodes & (n_nodes - 1)) == 0,
          "merkle_reduce_to_root: n_nodes must be a power of two <= 1024");
  hipLaunchKernelGGL(k_reduce_to_root, dim3(1), dim3(RTR_THREADS), 0, s, nodes, root, n_nodes);
  VX_HIP(hipGetLastError());
}

// Root of a tree of at most 2^MAX_TILE_LOG leaves in one workgroup, nothing
// stored (cuda_merkle_commit_small_soa4, cuda/blake2s.cu:92-189).
void merkle_root_small(const LeafSrc& L, uint32_t log_n, uint32_t* root_out, hipStream_t s) {
  check_leaf(L, log_n);
  require(log_n <= MAX_TILE_LOG && root_out, "merkle_root_small: at most 4096 leaves");
  TreeLevels T{};
  T.base = nullptr;
  T.log_n = log_n;
  RootHook hook{};
  hook.root_out = root_out;
  dispatch_subtree(L, T, log_n, lane_log_for(log_n), MAX_LEVELS, hook, s);
}

// parents[j] = H(leaf(2j), leaf(2j+1)) (cuda/blake2s.cu:195-240).
__global__ void k_leaf_pairs(LeafSrc L, uint64_t n_parents, uint32_t* parents) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_parents) return;
  uint32_t nc = L.n_cols > 16 ? 16 : L.n_cols;
  uint32_t h[2][8];
  for (int s = 0; s < 2; s++) {
    uint32_t w[16];
    for (uint32_t c = 0; c < 16; c++) w[c] = c < nc ? L.col[c][2 * j + s] : 0u;
    b2_hash_words(w, nc, h[s]);
  }
  uint32_t p[8];
  b2_hash_node(h[0], h[1], p);
  store_node(parents + 8 * j, p);
}
void merkle_leaf_pairs(const LeafSrc& L, uint64_t n_leaves, uint32_t* parents, hipStream_t s) {
  require(L.kind == LEAF_COLS && L.n_cols >= 1 && L.n_cols <= 16 && n_leaves % 2 == 0,
          "merkle_leaf_pairs: 1..16 columns, even leaf count");
  uint64_t np = n_leaves / 2;
  if (np == 0) return;
  hipLaunchKernelGGL(k_leaf_pairs, dim3((uint32_t)((np + 255) / 256)), dim3(256), 0, s, L, np, parents);
  VX_HIP(hipGetLastError());
}

__global__ void k_lifted_leaves(LeafSrc L, uint64_t n, uint32_t* hashes) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t h[8];
  lifted_leaf(L, i, h);
  store_node(hashes + 8 * i, h);
}
void merkle_lifted_leaves(const LeafSrc& L, uint32_t* hashes, hipStream_t s) {
  check_leaf(L, L.lift_log);
  const uint64_t n = 1ull << L.lift_log;
  hipLaunchKernelGGL(k_lifted_leaves, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, L, n, hashes);
  VX_HIP(hipGetLastError());
}

// ------------------------------------------------------------- decommitment
// One 64-lane workgroup per (tree, query, role): role 0 opens the query leaf,
// role 1 its partner (the sibling, or the next row). Each re-hashes the
// 2^lo-leaf group that holds its leaf, takes the path below level `lo` from
// it and the rest from the stored levels, and writes its values and auth
// path into the proof (src/bin/stark_cli.rs:1061-1089 layout: n, values,
// partner values, paths, partner paths).
__device__ __forceinline__ uint32_t decom_width(const DecomDesc& d) {
  return (d.kind == DECOM_COLS || d.kind == DECOM_LIFTED) ? d.n_cols : 4u;
}
__device__ __forceinline__ void decom_value(const DecomDesc& d, uint64_t i, uint32_t w[16]) {
  if (d.kind == DECOM_COLS) {
    for (uint32_t c = 0; c < d.n_cols; c++) w[c] = d.col[c][i];
  } else if (d.kind == DECOM_SOA4) {
    for (int c = 0; c < 4; c++) w[c] = d.col[c][i];
  } else {
    const uint64_t mask = (1ull << d.log_n) - 1;
    const uint32_t* ev = d.col[0];
    uint32_t c = m31_sub(ev[(i + 2) & mask], m31_add(ev[(i + 1) & mask], ev[i]));
    qm31 q = qm_scale(ldq(d.alpha), c);
    for (int k = 0; k < 4; k++) w[k] = q.v[k];
  }
}

__global__ __launch_bounds__(64) void k_decommit(const DecomDesc* descs, const uint32_t* queries, uint32_t nq,
                                                 uint32_t* proof) {
  __shared__ uint32_t nodes[32 * 8];  // level v of the group at node offset 32 - 2^(lo - v + 1)
  const uint32_t role = blockIdx.x & 1, tq = blockIdx.x >> 1;
  const uint32_t tree = tq / nq, q = tq % nq;
  const DecomDesc& d = descs[tree];
  const uint32_t lane = threadIdx.x;
  const uint32_t lo = d.lo, D = d.log_n, W = decom_width(d);
  const uint64_t mask = (1ull << D) - 1;
  const uint64_t qidx = (uint64_t)queries[q] >> d.shift;
  const uint64_t idx = role == 0 ? qidx : (d.pair == PAIR_NEXT ? (qidx + 1) & mask : qidx ^ 1);
  const uint64_t first = (idx >> lo) << lo;
  // level v occupies nodes [base(v), base(v) + 2^(lo-v)), base(v) = 32 - 2^(lo-v+1)
  if (lane < (1u << lo)) {
    uint32_t h[8];
    if (d.kind == DECOM_LIFTED) {
      lifted_leaf_cols(d.lift, d.n_cols, D, first + lane, h);
    } else {
      uint32_t w[16];
      for (int k = 0; k < 16; k++) w[k] = 0;
      decom_value(d, first + lane, w);
      b2_hash_words(w, W, h);
    }
    for (int k = 0; k < 8; k++) nodes[8 * (32 - (2u << lo) + lane) + k] = h[k];
  }
  __syncthreads();
  for (uint32_t v = 1; v < lo; v++) {
    const uint32_t cnt = 1u << (lo - v);
    const uint32_t src = 32 - (4u << (lo - v)), dst = 32 - (2u << (lo - v));
    if (lane < cnt) {
      uint32_t l[8], r[8], h[8];
      for (int k = 0; k < 8; k++) {
        l[k] = nodes[8 * (src + 2 * lane) + k];
        r[k] = nodes[8 * (src + 2 * lane + 1) + k];
      }
      b2_hash_node(l, r, h);
      for (int k = 0; k < 8; k++) nodes[8 * (dst + lane) + k] = h[k];
    }
    __syncthreads();
  }
  const uint64_t base = d.proof_word;
  const uint64_t pbase = base + 1 + 2ull * nq * W;
  const uint64_t plen = 1 + 8ull * D;
  uint32_t* path = proof + pbase + ((uint64_t)role * nq + q) * plen;
  if (d.kind == DECOM_LIFTED) {  // the opened row: every column at its lifted row
    for (uint32_t c = lane; c < W; c += 64)
      proof[base + 1 + ((uint64_t)role * nq + q) * W + c] = d.lift[c].p[lifted_row(idx, D, d.lift[c].log_n)];
    if (lane == 0) {
      if (q == 0 && role == 0) proof[base] = nq;
      path[0] = D;
    }
  } else if (lane == 0) {
    uint32_t w[16];
    if (q == 0 && role == 0) proof[base] = nq;
    decom_value(d, idx, w);
    for (uint32_t k = 0; k < W; k++) proof[base + 1 + ((uint64_t)role * nq + q) * W + k] = w[k];
    path[0] = D;
  }
  if (lane < D) {
    const uint32_t v = lane;
    const uint32_t* sib;
    if (v < lo) {
      const uint32_t lvl_base = 32 - (2u << (lo - v));
      sib = &nodes[8 * (lvl_base + (uint32_t)(((idx >> v) ^ 1) - (first >> v)))];
    } else {
      sib = d.tree + d.off[v] + 8 * ((idx >> v) ^ 1);
    }
    for (int k = 0; k < 8; k++) path[1 + 8 * v + k] = sib[k];
  }
}

void decommit(const DecomDesc* d_descs, uint32_t n_trees, const uint32_t* d_queries, uint32_t nq, uint32_t* proof,
              hipStream_t s) {
  if (n_trees == 0 || nq == 0) return;
  require((uint64_t)n_trees * nq * 2 < (1ull << 31), "decommit: too many openings");
  hipLaunchKernelGGL(k_decommit, dim3(2 * n_trees * nq), dim3(64), 0, s, d_descs, d_queries, nq, proof);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
