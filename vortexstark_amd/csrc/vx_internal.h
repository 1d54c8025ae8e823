// Internal interfaces between the gfx950 kernel translation units and the
// C++ host layer (prover, C-ABI shims). Not part of the public C-ABI.
#pragma once
#include <functional>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "vx_channel.h"
#include "vx_circle.h"
#include "vx_error.h"
#include "vx_field.h"

namespace vx {

// ---------------------------------------------------------------- errors
void hip_check(hipError_t e, const char* what);
#define VX_HIP(call) ::vx::hip_check((call), #call)
// Last error message of the calling thread (C-ABI vx_last_error).
void set_last_error(const std::string& msg);

// Read of wave-uniform data that no launch modifies while it runs (job
// lists, constraint and column tables): through the constant address space,
// so the compiler issues scalar (SMEM) loads and keeps the values in SGPRs
// instead of one vector load per lane, each a full memory round trip.
template <typename T>
using uniform_t = const __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ uniform_t<T>& as_uniform(const T* p) {
  return *(uniform_t<T>*)p;
}
// A pointer known to address global memory (e.g. one read from a table),
// so its loads are global rather than flat loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ uint4 ld4_global(const uint32_t* p) {
  const __attribute__((address_space(1))) uint32_t* q = as_global(p);
  return make_uint4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ uint2 ld2_global(const uint32_t* p) {
  const __attribute__((address_space(1))) uint32_t* q = as_global(p);
  return make_uint2(q[0], q[1]);
}
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "ld_uniform: whole words only");
  T r;
  uint32_t* d = reinterpret_cast<uint32_t*>(&r);
  uniform_t<uint32_t>* q = (uniform_t<uint32_t>*)p;
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; i++) d[i] = q[i];
  return r;
}

// --------------------------------------------------------- NTT (ntt.hip)
// Reference-layout twiddle tables of a coset (src/circle.rs:284-347):
//   line: layer l occupies [off[l], off[l] + (N >> (l+1))), circle: N/2 y-coords.
// stwo layout (src/ntt.rs:296-354, one flat buffer of N/2 words): line layer l
// at off[l] = N/2 - 2^(log_n - 1 - l), and stwo_circle set: the circle
// twiddle of butterfly t is derived from line layer 0 as
// (y, -y, -x, x)[t & 3] with (x, y) = (line[2(t >> 2)], line[2(t >> 2) + 1]),
// i.e. +-line[(t >> 1) ^ 1] (cuda/circle_ntt_stwo.cu:202-248); `circle` is unused.
struct TwiddleRef {
  const uint32_t* line;
  const uint32_t* circle;
  uint32_t off[32];
  uint32_t log_n;
  uint32_t stwo_circle;
};

// Multi-bit butterfly pass plan entry.
struct NttPass {
  uint32_t j0;     // lowest bit of the pass
  uint32_t k;      // number of bits
  uint32_t cbits;  // log2(columns per tile)
};

// Runs the twisted circle NTT over all log_n bits, in the reference layer
// order (cuda/circle_ntt.cu:248-312), using fused multi-bit LDS passes.
//   forward: evaluate; src (may alias data) is read with index mask
//            src_mask, which replicates a shorter coefficient vector into
//            every block (LDE, the first log_n - log(src) layers are copies).
//            Only bits below `eff_bits` are transformed.
//   inverse: interpolate, then scale by 2^((30 log_n) mod 31).
// Batched over n_cols columns (one per blockIdx.y): column c of data starts
// at data + c * dst_stride (stride for the inverse), of src at src + c * src_stride.
// Tunable schedule of a transform (ntt.hip): split = 1 spreads the strided
// bits over one more pass, order = 1 runs tiles in launch order instead of
// XCD-contiguous ranges. Every plan computes the same values.
struct NttPlan {
  uint32_t split = 0, order = 0;
};
struct NttKey {
  bool fwd;
  uint32_t log_n, eff_bits, n_cols;
  bool stwo;
};
// The plan of a transform shape (tune.cpp): cached, or measured now by
// calling run(plan) for each candidate (run returns false where it cannot
// time that shape: the built-in plan is used), or the built-in plan while the
// stream is being captured or tuning is off.
NttPlan tuned_ntt_plan(const NttKey& key, hipStream_t s, const std::function<bool(const NttPlan&)>& run);
NttPlan known_ntt_plan(const NttKey& key);  // cached choice or the built-in plan, never measures
constexpr uint64_t NTT_TUNE_SCRATCH_MAX_WORDS = 1ull << 28;

void ntt_forward(uint32_t* data, const uint32_t* src, uint64_t src_mask, const TwiddleRef& tw,
                 uint32_t eff_bits, hipStream_t s, uint32_t n_cols = 1, uint64_t dst_stride = 0,
                 uint64_t src_stride = 0);
void ntt_inverse(uint32_t* data, const TwiddleRef& tw, hipStream_t s, uint32_t n_cols = 1, uint64_t stride = 0);
// In-place transforms of n_cols columns of 2^log_n words each at arbitrary
// (16-byte aligned) addresses: d_cols is a device array of column pointers.
// Same passes as ntt_forward / ntt_inverse, one launch per pass for all columns.
// With d_src_cols, the forward transform reads 2^src_log_n coefficients per
// column from there, zero-padded (the LDE: the replication layers are skipped).
void ntt_forward_cols(uint32_t* const* d_cols, uint32_t n_cols, const TwiddleRef& tw, hipStream_t s,
                      const uint32_t* const* d_src_cols = nullptr, uint32_t src_log_n = 0);
void ntt_inverse_cols(uint32_t* const* d_cols, uint32_t n_cols, const TwiddleRef& tw, hipStream_t s);
// Uneven batched LDE (columns of log sizes 14..24): host-built work lists
// (lde_batch_build) uploaded by the caller, then four launches.
struct LdeBatchCol {
  uint32_t log_n;
  const uint32_t* in;  // 2^log_n trace values (null: read from coef, in place)
  uint32_t* coef;   // out: coefficients
  uint32_t* eval;   // out: 2^(log_n + blowup) evaluation
  TwiddleRef itw;   // inverse tables of half_coset(log_n)
  TwiddleRef etw;   // forward tables of the eval domain
};
struct LdeBatchStages {
  uint32_t count[4];
  uint64_t tiles[4];
  uint64_t offset[4];  // byte offset of each stage's jobs
};
size_t lde_batch_job_bytes();
void lde_batch_build(const LdeBatchCol* cols, uint32_t n, uint32_t blowup_bits, std::vector<uint8_t>& jobs,
                     LdeBatchStages& st);
void lde_batch_launch(const uint8_t* d_jobs, const LdeBatchStages& st, hipStream_t s);
// Out-of-place batched bit reverse of columns of log size >= 10.
void bitrev_batch_build(const uint32_t* const* src, uint32_t* const* dst, const uint32_t* log_n, uint32_t n,
                        std::vector<uint8_t>& jobs, uint32_t& count, uint64_t& tiles);
void bitrev_batch_launch(const uint8_t* d_jobs, uint32_t count, uint64_t tiles, hipStream_t s);
// One reference layer (cuda/circle_ntt.cu:89-117).
void ntt_layer(uint32_t* data, const uint32_t* tw, uint32_t layer_idx, uint64_t n, bool forward,
               hipStream_t s);
void m31_scale(uint32_t* data, uint32_t scale, uint64_t n, hipStream_t s);
void bit_reverse_m31(uint32_t* data, uint32_t log_n, hipStream_t s);
void bit_reverse_qm31(const uint32_t* in, uint32_t* out, uint32_t log_n, hipStream_t s);
void zero_pad(const uint32_t* src, uint32_t* dst, uint64_t src_n, uint64_t dst_n, hipStream_t s);

// ------------------------------------------- point evaluation (oods.hip)
// f(z) of 2^log_n stwo-basis coefficients from the QM31 folding factors
// (factors[0] folds the top coefficient bit). Uses s1 and s2 as scratch (each
// at least 4 words below 2^14 coefficients, else max(80, 2^log_n / 1024)
// words); returns the one holding the QM31.
const uint32_t* eval_at_point(const uint32_t* coeffs, const uint32_t* factors, uint32_t log_n, uint32_t* s1,
                              uint32_t* s2, hipStream_t s);
// Columns k < n_cols (HOST arrays of device pointers and log sizes <=
// max_log_n) at one point: results[k] (QM31, device) = f_k(z), column k
// using the last log_size_k of the max_log_n factors. Synchronous.
void eval_at_point_batch(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                         const uint32_t* factors, uint32_t max_log_n, uint32_t* results, hipStream_t s);
// The OODS step of src/oods.rs on a device channel (9 words: state, n_draws):
// draw z, z_next = z * step(trace_log); points = (z.x, z.y, z_next.x,
// z_next.y); values = f_k(z) for every column, then f_k(z_next); the 2 n_cols
// values mixed into the channel. Synchronous.
void oods_sample(uint32_t* d_channel, const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                 uint32_t trace_log, uint32_t* d_points, uint32_t* d_values, hipStream_t s);
// Per-block partial sums of evals[i] * weights[i] over the reference's
// grid-stride partition (n_blocks blocks of 256 lanes); out: n_blocks QM31.
void barycentric_partials(const uint32_t* evals, const uint32_t* weights, uint32_t n, uint32_t* out,
                          uint32_t n_blocks, hipStream_t s);

// ----------------------------------- stwo plugin QuotientOps (quotient.hip)
// Device pointer arrays, as the reference takes them. dual selects the fused
// z / z_next form (zn_idx, b_zn, c_zn used); otherwise only oz is written.
void accumulate_numerators(const uint32_t* const* cols, const uint32_t* idx_z, const uint32_t* zn_idx,
                           const uint32_t* b_z, const uint32_t* c_z, const uint32_t* b_zn, const uint32_t* c_zn,
                           uint32_t n_cols, uint32_t n_rows, bool dual, uint32_t* const oz[4],
                           uint32_t* const on[4],
                           hipStream_t s);
void quotients_combine(const uint32_t* sp_x, const uint32_t* sp_y, const uint32_t* fla,
                       const uint32_t* const* const numer[4], const uint32_t* acc_log_sizes, uint32_t n_accs,
                       const uint32_t* xs, const uint32_t* ys, uint32_t lifting_log_size, uint32_t n_rows,
                       uint32_t* const out[4], hipStream_t s);

// --------------------------------- stwo plugin column ops (column.hip)
void m31_binop(const uint32_t* a, const uint32_t* b, uint32_t* out, uint64_t n, bool mul, hipStream_t s);
void accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio, hipStream_t s);
// For each of 4 channels ch: last[ch][i] += sum_k src[ch * n_src + k][((i >> shift[k]) << 1) | (i & 1)]
// (device arrays: src pointers channel-major, shifts, the 4 last-column pointers).
void lift_accumulate(const uint32_t* const* d_src, const uint32_t* d_shift, uint32_t n_src, uint32_t* const* d_last,
                     uint32_t last_n, hipStream_t s);
// dst[i] = src[idx[i]] with `words` (1 or 8) words per element; checked skips
// elements whose source words are not all below src_len.
void gather(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, uint32_t words, bool checked,
            uint64_t src_len, hipStream_t s);
// 4 columns of n (a multiple of 16) words into 64 columns of n / 16:
// out[16 c + (p mod 16)][p / 16] = in[c][p] (device pointer arrays).
// Half-coset natural <-> canonic bit-reversed order; one permutation for
// both directions (it is an involution). Out of place.
void permute_hc_brt(const uint32_t* src, uint32_t* dst, uint32_t log_n, hipStream_t s);
void pack_leaves(const uint32_t* const* d_in, uint32_t* const* d_out, uint32_t n, hipStream_t s);

// ----------------------------------------------- twiddles (twiddles.hip)
void coset_points(cpt init, cpt step, uint32_t* xs, uint32_t* ys, uint64_t n, hipStream_t s);
void extract_and_squash(const uint32_t* in, uint32_t* tw_out, uint32_t* sq_out, uint64_t half_n,
                        hipStream_t s);
void batch_inverse(const uint32_t* in, uint32_t* out, uint64_t n, hipStream_t s);
void fold_twiddle_sources(cpt init, cpt step, uint32_t* out, uint64_t n, uint32_t log_n,
                          bool extract_y, hipStream_t s);
// stwo flat forward twiddles (n/2 words) of a 2^log_n-point transform and,
// if itw is set, their inverses (src/ntt.rs:296-354).
void stwo_twiddles(uint32_t log_n, uint32_t* tw, uint32_t* itw, hipStream_t s);

// ------------------------------------------------- misc kernels (fri.hip)
void fib_trace(uint32_t a, uint32_t b, uint32_t* out, uint32_t log_n, hipStream_t s);
void fib_trace_dev(const uint32_t* ab, uint32_t* out, uint32_t log_n, hipStream_t s);
void set_inputs(uint32_t* inputs, uint32_t* proof, uint32_t word_a, uint32_t a, uint32_t b, hipStream_t s);
void fib_quotient(const uint32_t* trace, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3,
                  qm31 alpha, uint64_t offset, uint64_t chunk_n, uint64_t global_n, hipStream_t s);
void fold_line(const uint32_t* const in[4], const uint32_t* tw, uint32_t* const out[4], qm31 alpha,
               uint64_t half_n, hipStream_t s);
void fold_circle_into_line(uint32_t* const dst[4], const uint32_t* const src[4], const uint32_t* tw,
                           qm31 alpha, qm31 alpha_sq, uint64_t half_n, hipStream_t s);

// --------------------------------------------------- Merkle (merkle.hip)
constexpr int MAX_LEVELS = 33;
// Stored levels [lo, log_n] of one Merkle tree; level l has 2^(log_n - l)
// nodes of 8 words at base + off[l].
struct TreeLevels {
  uint32_t* base;
  uint32_t log_n;
  uint32_t lo;
  uint64_t off[MAX_LEVELS];
};
// Arguments of the Cairo VM AIR quotient (vx_cairo.h).
struct CairoArgs;
// Leaf producers fused into the subtree kernel.
enum LeafKind : uint32_t {
  LEAF_COLS = 0,          // n_cols (1..16) M31 columns, leaf = Blake2s(row words)
  LEAF_QUOTIENT = 1,      // Fibonacci quotient of an eval column (cuda/constraint.cu:10-36)
  LEAF_FOLD_CIRCLE_Q = 2, // circle fold of that quotient (cuda/fri.cu:92-126), writes the layer
  LEAF_FOLD_LINE = 3,     // line fold of a QM31 SoA layer (cuda/fri.cu:58-87), writes the layer
  LEAF_D2_QUOTIENT = 4,   // synthetic AIR quotient of W eval columns (vx_d2.h), writes Q (SoA)
  LEAF_LIFTED = 5,        // lifted multi-size leaf (cuda/merkle_leaves_lifted.cu:1-98)
  LEAF_LIFTED_RLC = 6,    // random linear combination of the lifted rows, writes Q (SoA)
  LEAF_CAIRO_QUOTIENT = 7,  // Cairo VM AIR quotient of row i (cuda/cairo_constraint.cu), writes Q (SoA)
};
// A column of a lifted tree: 2^log_n values, log_n <= the tree's log size.
struct LiftCol {
  const uint32_t* p;
  uint32_t log_n;
  uint32_t pad;
};
struct LeafSrc {
  uint32_t kind;
  uint32_t n_cols;
  const uint32_t* col[16];   // LEAF_COLS: columns; QUOTIENT/FOLD_CIRCLE_Q: col[0] = eval;
                             // FOLD_LINE: col[0..3] = previous layer
  uint32_t* out[4];          // fold outputs (the new layer, SoA)
  const uint32_t* tw;        // fold twiddles
  const uint32_t* alpha_q;   // device QM31: quotient alpha
  const uint32_t* alpha_f;   // device QM31: fold alpha
  uint64_t src_mask;         // eval length - 1 (quotient wraparound)
  // LEAF_D2_QUOTIENT: col[0] = first of n_cols eval columns, col_stride
  // words apart; per constraint 8 words {c, d, a, b, e, 0, 0, 0}; beta =
  // per constraint 8 words {beta_k c_k, beta_k d_k} (QM31 each, d2_weights).
  // LEAF_LIFTED_RLC: beta = lift_n QM31 powers of the challenge
  uint64_t col_stride;
  uint32_t n_cons;
  const uint32_t* cons;
  const uint32_t* beta;
  // LEAF_LIFTED: lift_n columns sorted by ascending size, tree of 2^lift_log leaves
  const LiftCol* lift;
  uint32_t lift_n;
  uint32_t lift_log;
  // LEAF_CAIRO_QUOTIENT: the quotient's arguments in device memory (full
  // form: rows 0..2^log_n, next row wrapping); Q goes to out[0..3]
  const CairoArgs* cairo;
};
// What the kernel that produces a root does with it: copy it to root_out,
// and, with a channel, mix it (after ch_init when `init`) and draw the next
// alpha into alpha_out.
struct RootHook {
  Channel* ch;
  uint32_t* root_out;
  uint32_t* alpha_out;
  uint32_t init;
};
// Subtree schedule of one commitment: 2^tile_log leaves per workgroup and
// 2^lpt leaves per lane (lpt 0: one leaf per lane, tile levels hashed by
// quads; 2..4: a lane folds its leaves to one node in registers). Every
// valid shape produces the same tree; only the speed differs.
struct TreeShape {
  uint32_t tile_log, lpt;
};
// The built-in choice for a tree of 2^log_n leaves of this kind.
TreeShape merkle_default_shape(const LeafSrc& leaf, uint32_t log_n);
// The shapes merkle_commit accepts for this tree when levels from store_lo
// up are stored (store_lo between lpt and tile_log).
std::vector<TreeShape> merkle_shapes(uint32_t log_n, uint32_t store_lo);
// log2 of the leaves one subtree workgroup commits with the built-in shape
// (the lowest level merkle_commit can store above the lane level).
uint32_t merkle_tile_log(const LeafSrc& leaf, uint32_t log_n);
// Hashes all 2^log_n leaves and stores tree levels [store_lo, log_n] into
// `tl` (store_lo >= tl.lo, between the shape's lane level and tile size).
// The shape comes from the autotuner (tune.cpp) when it is enabled, else
// the built-in choice. Optional events bracket the subtree kernel.
void merkle_commit(const LeafSrc& leaf, const TreeLevels& tl, uint32_t store_lo, const RootHook& hook,
                   hipStream_t s, hipEvent_t sub_begin = nullptr, hipEvent_t sub_end = nullptr);
// The same with an explicit shape (validated).
void merkle_commit_shaped(const LeafSrc& leaf, const TreeLevels& tl, uint32_t store_lo, const RootHook& hook,
                          hipStream_t s, const TreeShape& shape, hipEvent_t sub_begin = nullptr,
                          hipEvent_t sub_end = nullptr);
// Autotuner (tune.cpp): the fastest shape for this commitment, measured on
// the stream the first time its key (leaf kind, column / constraint counts,
// log_n, store_lo) is seen, and cached for the process (and in the cache
// file when one is configured). Never measures while the stream is being
// captured into a graph; then the cached or built-in shape is returned.
TreeShape tuned_shape(const LeafSrc& leaf, const TreeLevels& tl, uint32_t store_lo, hipStream_t s);
// FRI line-fold layers of <= 2^TAIL_MAX_LOG values, committed by one
// workgroup in one launch.
constexpr uint32_t TAIL_MAX_LOG = 10;
constexpr uint32_t TAIL_MAX_LAYERS = TAIL_MAX_LOG;
struct FriTailLayer {
  uint32_t* out[4];            // layer values (SoA), 2^log_n each
  const uint32_t* tw;          // fold twiddles of the fold that produces this layer
  uint32_t* tree;              // stored tree levels [lo, log_n]
  uint32_t off[TAIL_MAX_LOG + 1];
  uint32_t* root_out;          // proof slot of the root
  const uint32_t* alpha;       // device QM31 used by this fold
  uint32_t* next_alpha;        // where to draw the next fold's alpha (null: none)
  uint32_t log_n, lo;
};
struct FriTailArgs {
  const uint32_t* in[4];       // layer folded into the first tail layer
  Channel* ch;
  uint32_t n_layers;
  FriTailLayer layer[TAIL_MAX_LAYERS];
};
void fri_tail(const FriTailArgs& a, hipStream_t s);
// Tiled commit (cuda/blake2s.cu:305-532): roots of each 2^tile_log-leaf tile.
void merkle_tiled(const LeafSrc& leaf, uint32_t log_n, uint32_t tile_log, uint32_t* subtree_roots,
                  hipStream_t s);
void merkle_leaf_hashes(const LeafSrc& leaf, uint64_t n_leaves, uint32_t* hashes, hipStream_t s);
void merkle_hash_nodes(const uint32_t* children, uint32_t* parents, uint64_t n_parents, hipStream_t s);
void merkle_reduce_to_root(const uint32_t* nodes, uint32_t* root, uint32_t n_nodes, hipStream_t s);
void merkle_root_small(const LeafSrc& leaf, uint32_t log_n, uint32_t* root_out, hipStream_t s);
void merkle_leaf_pairs(const LeafSrc& leaf, uint64_t n_leaves, uint32_t* parents, hipStream_t s);
// Every leaf hash of a lifted tree (cuda_build_leaves_lifted semantics).
void merkle_lifted_leaves(const LeafSrc& leaf, uint32_t* hashes, hipStream_t s);
// Cairo VM AIR quotient (cairo.hip).
void cairo_quotient(const CairoArgs& a, uint64_t rows, hipStream_t s);
void cairo_vanishing(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* out, uint32_t log_eval,
                     uint32_t log_n, bool inverse, hipStream_t s);
void axpy_m31(uint32_t sc, const uint32_t* x, uint32_t* y, uint64_t n, hipStream_t s);

// ------------------------------------------- bytecode constraint eval (bytecode.hip)
// Evaluates a program renamed by bc_compile (d_ins: its 16-byte records,
// d_consts: its QM31 constants, d_trace: its (column, offset) pairs) over
// n_rows rows; d_ws holds bytecode_ws_bytes(p, n_rows) bytes (null when
// that is 0).
struct BcProgram;
void bytecode_constraint_eval(const BcProgram& p, const uint32_t* d_ins, const uint32_t* d_consts,
                              const uint32_t* d_trace, const uint32_t* const* d_cols, uint32_t n_rows, uint32_t trace_n_rows,
                              const uint32_t* coeff, const uint32_t* denom_inv, uint32_t* const acc[4],
                              uint32_t* d_ws, hipStream_t s);
uint32_t bytecode_row_words(const BcProgram& p);
// The compiled tier (bytecode_jit.hip): the program (host words, checked by
// bc_compile) compiled once per device and domain shape, then launched.
void bytecode_jit_eval(const uint32_t* words, uint32_t n_words, const uint32_t* const* d_cols, uint32_t n_trace_cols,
                       uint32_t n_rows, uint32_t trace_n_rows, const uint32_t* coeff, const uint32_t* denom_inv,
                       uint32_t* const acc[4], hipStream_t s);
void bytecode_jit_enable(bool on);
// Generates and compiles (no device needed) and returns the code object's size.
size_t bytecode_jit_code_bytes(const uint32_t* words, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                               uint32_t trace_log);
bool bytecode_jit_enabled();
uint32_t bytecode_ws_rows(const BcProgram& p, uint32_t n_rows);
uint64_t bytecode_ws_bytes(const BcProgram& p, uint32_t n_rows);
// Root of a lifted tree with the column table and the stored tree levels in
// a file-scope device scratch array (noalloc.hip): no device allocation.
// Asynchronous; calls that overlap on one GPU share the scratch and race.
void merkle_root_lifted_noalloc(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                                uint32_t* d_root, hipStream_t s);

// Decommitment of one tree at the query positions (src/prover.rs:551-618),
// written straight into the KRAK byte buffer.
// Values of a leaf: DECOM_COLS = n_cols (<= 16) M31 columns, DECOM_SOA4 = a
// QM31 SoA layer, DECOM_QUOTIENT = the Fibonacci quotient recomputed from
// the eval column. The partner leaf of query leaf i is i ^ 1 (PAIR_SIBLING,
// the KRAK layout) or (i + 1) mod 2^log_n (PAIR_NEXT, next-row openings).
enum DecomKind : uint32_t { DECOM_COLS = 0, DECOM_SOA4 = 1, DECOM_QUOTIENT = 2, DECOM_LIFTED = 3 };
enum DecomPair : uint32_t { PAIR_SIBLING = 0, PAIR_NEXT = 1 };
struct DecomDesc {
  uint32_t kind;
  uint32_t n_cols;       // DECOM_COLS
  uint32_t pair;
  uint32_t log_n;
  uint32_t lo;           // lowest stored tree level
  uint32_t shift;        // leaf index = query >> shift
  const uint32_t* col[16];
  const uint32_t* alpha; // DECOM_QUOTIENT
  const LiftCol* lift;   // DECOM_LIFTED: n_cols columns, tree of 2^log_n leaves
  const uint32_t* tree;
  uint64_t off[MAX_LEVELS];
  uint64_t proof_word;   // word offset of this decommitment ("n") in the proof
};
void decommit(const DecomDesc* d_descs, uint32_t n_trees, const uint32_t* d_queries, uint32_t nq,
              uint32_t* proof, hipStream_t s);

// ------------------------------------------------- GKR (gkr.hip)
// The stwo plugin's GKR operations (cuda/gkr.cu); kinds 0 grand product,
// 1 LogUp generic, 2 LogUp multiplicities, 3 LogUp singles.
void gkr_fix_first_variable(const uint32_t* in, uint32_t* out, const uint32_t* r, uint64_t n, bool secure,
                            hipStream_t s);
void gkr_gen_eq_evals_init(uint32_t* buf, const uint32_t* v, hipStream_t s);
void gkr_gen_eq_evals_pass(uint32_t* buf, const uint32_t* y, uint64_t cur, hipStream_t s);
void gkr_gen_eq_evals(uint32_t* buf, const uint32_t* y, uint32_t k, const uint32_t* v, hipStream_t s);
void gkr_next_layer(int kind, const uint32_t* num, const uint32_t* den, uint32_t* out_num, uint32_t* out_den,
                    uint64_t n, hipStream_t s);
// Returns the number of 8-word blocks written to partial (1).
uint32_t gkr_sum_poly(int kind, const uint32_t* eq, const uint32_t* num, const uint32_t* den, const uint32_t* lambda,
                      uint32_t* partial, uint64_t n_terms, hipStream_t s);

// ------------------------------------------------- channel (channel.hip)
// Single-lane transcript steps over a device-resident Channel.
void channel_step(Channel* ch, bool init, const uint32_t* digest, uint32_t* alpha_out,
                  uint32_t* alpha_sq_out, hipStream_t s);
struct QueryArgs {
  Channel* ch;
  const uint32_t* last[4];  // last FRI layer, SoA, 8 values
  uint32_t pow_bits;
  const uint64_t* nonce;    // device, when pow_bits > 0
  uint32_t nq;
  uint64_t domain;          // eval domain size
  uint32_t* queries;        // out, nq indices
  uint32_t* proof;          // KRAK buffer
  uint64_t last_word;       // word offset of n_last
  uint64_t query_word;      // word offset of n_queries
  uint64_t nonce_word;      // word offset of nonce (v2) or 0
};
void channel_mix_last(const QueryArgs& q, uint32_t* prefix_out, hipStream_t s);
void channel_draw_queries(const QueryArgs& q, hipStream_t s);
// Deterministic PoW: smallest nonce >= 0 (linear order) whose Blake2s has
// >= pow_bits trailing zeros (src/channel.rs:140-168).
void grind_pow(const uint32_t* d_prefix, uint32_t pow_bits, uint64_t* d_result, hipStream_t s);
// The same search over the stwo plugin's nonces (hi << 32) | low, low < 2^20;
// m31_out reduces the hash words mod P before the count (stwo's
// Blake2sChannelGeneric<true>).
void grind_pow_stwo(const uint32_t* d_prefix, uint32_t pow_bits, uint64_t* d_result, bool m31_out, hipStream_t s);
void grind_pow_batch(const uint32_t* d_prefix, uint64_t* d_result, uint32_t pow_bits,
                     uint64_t batch_offset, uint64_t n_nonces, bool m31_out, hipStream_t s);
// Poseidon252 (Stark field, Montgomery form 2^256, elements as 4 x u64):
// the grind over nonces [batch_offset, batch_offset + n) lowering *d_result
// to the smallest valid one; the whole search from nonce 0; lifted leaf
// hashes; parent layers.
void poseidon_grind_batch(const uint64_t* d_prefix_mont, uint64_t* d_result, uint32_t pow_bits, uint64_t batch_offset,
                          uint64_t n, hipStream_t s);
void poseidon_grind_min(const uint64_t* d_prefix_mont, uint32_t pow_bits, uint64_t* d_result, hipStream_t s);
void poseidon_leaves(const uint32_t* const* d_cols, const uint32_t* d_logs, uint32_t n_cols, uint32_t lifting_log,
                     uint64_t* d_out, uint32_t n_leaves, hipStream_t s);
void poseidon_next_layer(const uint64_t* d_prev, uint64_t* d_out, uint32_t n_parents, hipStream_t s);
void write_header(uint32_t* proof, const uint32_t* words, uint32_t n, hipStream_t s);
// Copy `words` words from device memory into pinned host memory
// (hipHostMalloc) with a kernel; both 16-byte aligned.
void copy_to_host(uint32_t* host_dst, const uint32_t* dev_src, uint64_t words, hipStream_t s);

// ------------------------------------- synthetic multi-column AIR (d2.hip)
void d2_trace(const uint32_t* seed_words, uint32_t* out, uint32_t n_cols, uint32_t log_n, uint64_t stride,
              hipStream_t s);
void d2_powers(const uint32_t* alpha, uint32_t* beta, uint32_t k_cons, hipStream_t s);
void d2_weights(const uint32_t* alpha, const uint32_t* cons, uint32_t* w, uint32_t k_cons, hipStream_t s);
void d2_trace_col(const uint32_t* seed_words, uint32_t* out, uint32_t col, uint32_t log_n, hipStream_t s);
void d2_fill_columns(const uint32_t* seed_words, uint32_t* out, const uint64_t* d_off, uint32_t n_cols,
                     uint64_t total, hipStream_t s);

}  // namespace vx
