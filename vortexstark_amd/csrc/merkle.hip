// Blake2s Merkle commitments for gfx950.
//
// Tree semantics (cuda/blake2s.cu:13-87, src/merkle.rs:520-593,841-884):
// leaf = Blake2s(row words, t = 4 * n_cols), node = Blake2s(left || right,
// t = 64), leaves in natural order, parent j = H(2j, 2j+1).
//
// MI355X design. One workgroup commits a 4096-leaf tile: each of 256 lanes
// hashes 16 consecutive leaves and folds them to a level-4 node entirely in
// VGPRs (31 compressions, state + message in registers), then levels 5..12
// are merged through LDS. Only levels >= 4 are written to HBM (about 4 bytes
// per leaf), which is all the decommitment needs: it re-hashes the 16-leaf
// group around each query. The leaf words can be produced on the fly by the
// fused producers below, so the Fibonacci quotient is never materialised
// and each FRI fold writes its layer once while hashing it:
//   LEAF_QUOTIENT       q_i = alpha (e[i+2] - e[i+1] - e[i])  (cuda/constraint.cu:10-36)
//   LEAF_FOLD_CIRCLE_Q  circle fold of that quotient           (cuda/fri.cu:92-126)
//   LEAF_FOLD_LINE      line fold of the previous layer        (cuda/fri.cu:58-87)
// Blake2s is VALU-bound here; loads are 16-byte vectors and hide under it.
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t SUB_THREADS = 256;
static constexpr uint32_t SUB_LPT_LOG = 4;   // leaves per lane (log)
static constexpr uint32_t SUB_TILE_LOG = 12; // leaves per workgroup (log)
static constexpr uint32_t RED_NLEV_MAX = 9;  // tree levels per reduce workgroup

struct G4 {
  uint32_t w[4][4];  // words of 4 consecutive leaves (W <= 4)
};

__device__ __forceinline__ uint4 ld4(const uint32_t* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ void st4(uint32_t* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  *reinterpret_cast<uint4*>(p) = make_uint4(a, b, c, d);
}
__device__ __forceinline__ qm31 ldq(const uint32_t* p) { return {{p[0], p[1], p[2], p[3]}}; }

// Fibonacci constraint for 8 consecutive positions from e[0..10).
__device__ __forceinline__ uint32_t fib_c(const uint32_t* e, int k) {
  return m31_sub(e[k + 2], m31_add(e[k + 1], e[k]));
}

// Produce the words of leaves [i, i+4) (and write fold outputs).
template <uint32_t KIND, uint32_t NC>
__device__ __forceinline__ void leaf_group(const LeafSrc& L, uint64_t i, G4& g) {
  if constexpr (KIND == LEAF_COLS && NC == 1) {
    uint4 v = ld4(L.col[0] + i);
    g.w[0][0] = v.x;
    g.w[1][0] = v.y;
    g.w[2][0] = v.z;
    g.w[3][0] = v.w;
  } else if constexpr (KIND == LEAF_COLS && NC == 4) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      uint4 v = ld4(L.col[c] + i);
      g.w[0][c] = v.x;
      g.w[1][c] = v.y;
      g.w[2][c] = v.z;
      g.w[3][c] = v.w;
    }
  } else if constexpr (KIND == LEAF_QUOTIENT) {
    const uint32_t* ev = L.col[0];
    uint4 v = ld4(ev + i);
    uint32_t e[6] = {v.x, v.y, v.z, v.w, ev[(i + 4) & L.src_mask], ev[(i + 5) & L.src_mask]};
    qm31 a = ldq(L.alpha_q);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qm31 q = qm_scale(a, fib_c(e, j));
#pragma unroll
      for (int c = 0; c < 4; c++) g.w[j][c] = q.v[c];
    }
  } else if constexpr (KIND == LEAF_FOLD_CIRCLE_Q) {
    const uint32_t* ev = L.col[0];
    const uint64_t b = 2 * i;
    uint4 v0 = ld4(ev + b), v1 = ld4(ev + b + 4);
    uint32_t e[10] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, ev[(b + 8) & L.src_mask],
                      ev[(b + 9) & L.src_mask]};
    qm31 aq = ldq(L.alpha_q), af = ldq(L.alpha_f);
    uint4 tw = ld4(L.tw + i);
    uint32_t t[4] = {tw.x, tw.y, tw.z, tw.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qm31 f0 = qm_scale(aq, fib_c(e, 2 * j)), f1 = qm_scale(aq, fib_c(e, 2 * j + 1));
      qm31 r = fri_fold_pair(f0, f1, af, t[j]);
#pragma unroll
      for (int c = 0; c < 4; c++) g.w[j][c] = r.v[c];
    }
#pragma unroll
    for (int c = 0; c < 4; c++) st4(L.out[c] + i, g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
  } else if constexpr (KIND == LEAF_FOLD_LINE) {
    const uint64_t b = 2 * i;
    uint32_t in[4][8];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      uint4 v0 = ld4(L.col[c] + b), v1 = ld4(L.col[c] + b + 4);
      in[c][0] = v0.x; in[c][1] = v0.y; in[c][2] = v0.z; in[c][3] = v0.w;
      in[c][4] = v1.x; in[c][5] = v1.y; in[c][6] = v1.z; in[c][7] = v1.w;
    }
    qm31 af = ldq(L.alpha_f);
    uint4 tw = ld4(L.tw + i);
    uint32_t t[4] = {tw.x, tw.y, tw.z, tw.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      qm31 f0 = {{in[0][2 * j], in[1][2 * j], in[2][2 * j], in[3][2 * j]}};
      qm31 f1 = {{in[0][2 * j + 1], in[1][2 * j + 1], in[2][2 * j + 1], in[3][2 * j + 1]}};
      qm31 r = fri_fold_pair(f0, f1, af, t[j]);
#pragma unroll
      for (int c = 0; c < 4; c++) g.w[j][c] = r.v[c];
    }
#pragma unroll
    for (int c = 0; c < 4; c++) st4(L.out[c] + i, g.w[0][c], g.w[1][c], g.w[2][c], g.w[3][c]);
  }
}

template <uint32_t KIND, uint32_t NC>
__device__ __forceinline__ void leaf_hash(const LeafSrc& L, const G4& g, uint64_t i, int j, uint32_t h[8]) {
  if constexpr (KIND == LEAF_COLS && NC == 1) {
    b2_leaf1(g.w[j][0], h);
  } else if constexpr (KIND == LEAF_COLS && NC == 0) {
    // generic 1..16 columns, read per leaf
    uint32_t w[16];
#pragma unroll
    for (uint32_t c = 0; c < 16; c++) w[c] = c < L.n_cols ? L.col[c][i + j] : 0u;
    b2_hash_words(w, L.n_cols > 16 ? 16 : L.n_cols, h);
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
  uint32_t s2[8], s3[8];
#pragma unroll
  for (uint32_t gi = 0; gi < (1u << (LPT_LOG - 2)); gi++) {
    const uint64_t i = first + 4 * gi;
    G4 g;
    if constexpr (!(KIND == LEAF_COLS && NC == 0)) leaf_group<KIND, NC>(L, i, g);
    uint32_t a[8], b[8], n1a[8], n1b[8], n2[8];
    leaf_hash<KIND, NC>(L, g, i, 0, a);
    leaf_hash<KIND, NC>(L, g, i, 1, b);
    b2_hash_node(a, b, n1a);
    leaf_hash<KIND, NC>(L, g, i, 2, a);
    leaf_hash<KIND, NC>(L, g, i, 3, b);
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

// Merges `nthr` nodes held one per lane in `buf[0]` up to one node, storing
// each produced level (lvl0+1 ...) at global index base_idx(level) + lane.
// Returns the top node in lane 0's `node`.
template <uint32_t CAP>
__device__ __forceinline__ void lds_merge(uint4 (*buf)[2 * CAP], uint32_t nthr, uint32_t lvl0,
                                          const TreeLevels& T, uint64_t group, uint32_t group_log,
                                          uint32_t store_lo, uint32_t node[8]) {
  const uint32_t tid = threadIdx.x;
  uint32_t cur = 0;
  uint32_t cnt = nthr;
  for (uint32_t lvl = lvl0 + 1; cnt > 1; lvl++) {
    cnt >>= 1;
    if (tid < cnt) {
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
}

template <uint32_t KIND, uint32_t NC, uint32_t LPT_LOG>
__global__ __launch_bounds__(SUB_THREADS) void k_subtree(LeafSrc L, TreeLevels T, uint32_t tile_log,
                                                         uint32_t store_lo, uint32_t* root_out) {
  __shared__ uint4 buf[2][2 * SUB_THREADS];
  const uint32_t nthr = 1u << (tile_log - LPT_LOG);
  const uint32_t tid = threadIdx.x;
  const uint64_t tile = blockIdx.x;
  uint32_t node[8];
  if (tid < nthr) {
    const uint64_t first = (tile << tile_log) + ((uint64_t)tid << LPT_LOG);
    lane_subtree<KIND, NC, LPT_LOG>(L, first, node);
    lds_put(buf[0], tid, node);
    if (T.base && LPT_LOG >= store_lo)
      store_node(T.base + T.off[LPT_LOG] + 8 * ((tile << (tile_log - LPT_LOG)) + tid), node);
  }
  __syncthreads();
  lds_merge<SUB_THREADS>(buf, nthr, LPT_LOG, T, tile, tile_log, store_lo, node);
  if (root_out && tid == 0 && tile_log == T.log_n) {
    for (int k = 0; k < 8; k++) root_out[k] = node[k];
  }
}

// Upper tree levels: each workgroup hashes 2^nlev level-`from` nodes down to
// one node, storing levels from+1 .. from+nlev.
__global__ __launch_bounds__(SUB_THREADS) void k_tree_reduce(TreeLevels T, uint32_t from, uint32_t nlev,
                                                             uint32_t* root_out) {
  __shared__ uint4 buf[2][2 * SUB_THREADS];
  const uint32_t tid = threadIdx.x;
  const uint32_t nthr = 1u << (nlev - 1);
  const uint64_t wg = blockIdx.x;
  uint32_t node[8];
  if (tid < nthr) {
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
  lds_merge<SUB_THREADS>(buf, nthr, from + 1, T, wg, from + nlev, 0, node);
  if (root_out && tid == 0 && from + nlev == T.log_n) {
    for (int k = 0; k < 8; k++) root_out[k] = node[k];
  }
}

template <uint32_t KIND, uint32_t NC>
static void launch_subtree(const LeafSrc& L, const TreeLevels& T, uint32_t tile_log, uint32_t store_lo,
                           uint32_t* root_out, hipStream_t s) {
  const uint64_t grid = 1ull << (T.log_n - tile_log);
  const uint32_t lpt = tile_log < SUB_LPT_LOG ? tile_log : SUB_LPT_LOG;
  require(lpt >= 3 || (KIND == LEAF_COLS && lpt >= 2), "merkle: tree too small");
  if (lpt == 4)
    hipLaunchKernelGGL((k_subtree<KIND, NC, 4>), dim3((uint32_t)grid), dim3(SUB_THREADS), 0, s, L, T, tile_log,
                       store_lo, root_out);
  else if (lpt == 3)
    hipLaunchKernelGGL((k_subtree<KIND, NC, 3>), dim3((uint32_t)grid), dim3(SUB_THREADS), 0, s, L, T, tile_log,
                       store_lo, root_out);
  else
    hipLaunchKernelGGL((k_subtree<KIND, NC, 2>), dim3((uint32_t)grid), dim3(SUB_THREADS), 0, s, L, T, tile_log,
                       store_lo, root_out);
  VX_HIP(hipGetLastError());
}

static void check_leaf(const LeafSrc& L, uint32_t log_n) {
  require(log_n >= 2 && log_n <= 31, "merkle: log_n out of range");
  switch (L.kind) {
    case LEAF_COLS:
      require(L.n_cols >= 1 && L.n_cols <= 16, "merkle: n_cols must be 1..16");
      for (uint32_t c = 0; c < L.n_cols; c++) require(L.col[c] != nullptr, "merkle: null column");
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
    default:
      require(false, "merkle: unknown leaf kind");
  }
}

static void dispatch_subtree(const LeafSrc& L, const TreeLevels& T, uint32_t tile_log, uint32_t store_lo,
                             uint32_t* root_out, hipStream_t s) {
  switch (L.kind) {
    case LEAF_COLS:
      if (L.n_cols == 1) launch_subtree<LEAF_COLS, 1>(L, T, tile_log, store_lo, root_out, s);
      else if (L.n_cols == 4) launch_subtree<LEAF_COLS, 4>(L, T, tile_log, store_lo, root_out, s);
      else launch_subtree<LEAF_COLS, 0>(L, T, tile_log, store_lo, root_out, s);
      break;
    case LEAF_QUOTIENT: launch_subtree<LEAF_QUOTIENT, 4>(L, T, tile_log, store_lo, root_out, s); break;
    case LEAF_FOLD_CIRCLE_Q: launch_subtree<LEAF_FOLD_CIRCLE_Q, 4>(L, T, tile_log, store_lo, root_out, s); break;
    case LEAF_FOLD_LINE: launch_subtree<LEAF_FOLD_LINE, 4>(L, T, tile_log, store_lo, root_out, s); break;
  }
}

void merkle_commit(const LeafSrc& L, const TreeLevels& T, uint32_t store_lo, uint32_t* root_out, hipStream_t s,
                   hipEvent_t sub_begin, hipEvent_t sub_end) {
  check_leaf(L, T.log_n);
  require(T.base != nullptr, "merkle_commit: null tree buffer");
  const uint32_t tile_log = T.log_n < SUB_TILE_LOG ? T.log_n : SUB_TILE_LOG;
  const uint32_t lpt = tile_log < SUB_LPT_LOG ? tile_log : SUB_LPT_LOG;
  require(T.lo <= lpt && store_lo >= T.lo && store_lo <= lpt, "merkle_commit: stored levels must start at the lane level");
  if (sub_begin) VX_HIP(hipEventRecord(sub_begin, s));
  dispatch_subtree(L, T, tile_log, store_lo, root_out, s);
  if (sub_end) VX_HIP(hipEventRecord(sub_end, s));
  uint32_t from = tile_log;
  while (from < T.log_n) {
    uint32_t nlev = T.log_n - from < RED_NLEV_MAX ? T.log_n - from : RED_NLEV_MAX;
    uint64_t grid = 1ull << (T.log_n - from - nlev);
    hipLaunchKernelGGL(k_tree_reduce, dim3((uint32_t)grid), dim3(SUB_THREADS), 0, s, T, from, nlev, root_out);
    VX_HIP(hipGetLastError());
    from += nlev;
  }
}

void merkle_tiled(const LeafSrc& L, uint32_t log_n, uint32_t tile_log, uint32_t* subtree_roots, hipStream_t s) {
  check_leaf(L, log_n);
  require(tile_log >= 4 && tile_log <= SUB_TILE_LOG && tile_log <= log_n, "merkle_tiled: tile size out of range");
  TreeLevels T{};
  T.base = subtree_roots;
  T.log_n = log_n;
  T.lo = tile_log;
  T.off[tile_log] = 0;
  dispatch_subtree(L, T, tile_log, tile_log, nullptr, s);
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
  lds_merge<RTR_THREADS>(buf, nthr, 0, none, 0, 0, 0, node);
  if (tid == 0)
    for (int k = 0; k < 8; k++) root[k] = node[k];
}
void merkle_reduce_to_root(const uint32_t* nodes, uint32_t* root, uint32_t n_nodes, hipStream_t s) {
  require(n_nodes >= 1 && n_nodes <= 2 * RTR_THREADS && (n_nodes & (n_nodes - 1)) == 0,
          "merkle_reduce_to_root: n_nodes must be a power of two <= 1024");
  hipLaunchKernelGGL(k_reduce_to_root, dim3(1), dim3(RTR_THREADS), 0, s, nodes, root, n_nodes);
  VX_HIP(hipGetLastError());
}

// Root of a tree of at most 2^SUB_TILE_LOG leaves in one workgroup, nothing
// stored (cuda_merkle_commit_small_soa4, cuda/blake2s.cu:92-189).
void merkle_root_small(const LeafSrc& L, uint32_t log_n, uint32_t* root_out, hipStream_t s) {
  check_leaf(L, log_n);
  require(log_n <= SUB_TILE_LOG && root_out, "merkle_root_small: at most 4096 leaves");
  TreeLevels T{};
  T.base = nullptr;
  T.log_n = log_n;
  dispatch_subtree(L, T, log_n, MAX_LEVELS, root_out, s);
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

// ------------------------------------------------------------- decommitment
// One 64-lane workgroup per (tree, query): re-hash the 2^lo-leaf group that
// holds the query, take siblings below level `lo` from it and the rest from
// the stored levels, and write values, sibling values and both auth paths
// into the KRAK buffer (src/bin/stark_cli.rs:1061-1089 layout).
__device__ __forceinline__ void decom_value(const DecomDesc& d, uint64_t i, uint32_t w[4]) {
  if (d.kind == DECOM_M31) {
    w[0] = d.col[0][i];
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
  const uint32_t tree = blockIdx.x / nq, q = blockIdx.x % nq;
  const DecomDesc& d = descs[tree];
  const uint32_t lane = threadIdx.x;
  const uint32_t lo = d.lo, D = d.log_n, W = d.kind == DECOM_M31 ? 1u : 4u;
  const uint64_t idx = (uint64_t)queries[q] >> d.shift;
  const uint64_t first = (idx >> lo) << lo;
  // level v occupies nodes [base(v), base(v) + 2^(lo-v)), base(v) = 32 - 2^(lo-v+1)
  if (lane < (1u << lo)) {
    uint32_t w[4] = {0, 0, 0, 0}, h[8];
    decom_value(d, first + lane, w);
    b2_hash_words(w, W, h);
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
  uint32_t* path = proof + pbase + q * plen;
  uint32_t* spath = proof + pbase + (uint64_t)nq * plen + q * plen;
  if (lane == 0) {
    uint32_t w[4] = {0, 0, 0, 0};
    if (q == 0) proof[base] = nq;
    decom_value(d, idx, w);
    for (uint32_t k = 0; k < W; k++) proof[base + 1 + (uint64_t)q * W + k] = w[k];
    decom_value(d, idx ^ 1, w);
    for (uint32_t k = 0; k < W; k++) proof[base + 1 + (uint64_t)nq * W + (uint64_t)q * W + k] = w[k];
    path[0] = D;
    spath[0] = D;
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
    const uint32_t* ssib = sib;
    if (v == 0) ssib = &nodes[8 * (32 - (2u << lo) + (uint32_t)(idx - first))];
    for (int k = 0; k < 8; k++) spath[1 + 8 * v + k] = ssib[k];
  }
}

void decommit(const DecomDesc* d_descs, uint32_t n_trees, const uint32_t* d_queries, uint32_t nq, uint32_t* proof,
              hipStream_t s) {
  if (n_trees == 0 || nq == 0) return;
  hipLaunchKernelGGL(k_decommit, dim3(n_trees * nq), dim3(64), 0, s, d_descs, d_queries, nq, proof);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
