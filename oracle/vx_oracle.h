/*
 * vx_oracle.h -- CPU restatement of the VortexSTARK Fibonacci Circle-STARK
 * prove path, used ONLY as the correctness oracle and CPU baseline.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or the
 * timed CPU baseline). The product (vortexstark_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates. All arithmetic
 * is exact over GF(2^31-1), so results are bit-exact by construction.
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   field / circle        reference literal KATs, Python big-int identities
 *   blake2s / channel     Python hashlib.blake2s KATs (standard RFC 7693)
 *   twiddles / NTT        independent Python big-int restatement fixtures,
 *                         closed form at n=4, roundtrip, layer equivalence,
 *                         LDE replication property
 *   quotient / folds      closed form per element (Python big-int)
 *   merkle                Python hashlib tree, auth-path verification
 *   prove / verify        independent pure-Python prover byte-equality
 *                         (log_n 4..6), verifier accept + tamper-reject
 *   pow                   brute-force smallest nonce via hashlib
 *   d2 AIR (build-defined) independent Python restatement (pyref.prove_d2),
 *                         SplitMix64 reference output, closed forms
 *   lifted Merkle         Python hashlib over the lifted rows; the LDE is the
 *                         pinned single-column LDE per column
 *   cairo quotient        independent Python restatement (tests/golden/
 *                         pyref_cairo.py) of the reference's host-side
 *                         evaluation; valid Cairo rows give zero, one changed
 *                         value gives nonzero; vanishing at the trace domain
 *   ntt_evaluate_at       the full transform (or_ntt_evaluate, or_lde) at
 *                         every output for small sizes and three coset kinds,
 *                         and the roundtrip to the trace values
 */
#ifndef VX_ORACLE_H
#define VX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P 0x7FFFFFFFu

/* ---- field (src/field/m31.rs, cm31.rs, qm31.rs) ---- */
uint32_t or_m31_add(uint32_t a, uint32_t b);
uint32_t or_m31_sub(uint32_t a, uint32_t b);
uint32_t or_m31_mul(uint32_t a, uint32_t b);
uint32_t or_m31_inv(uint32_t a);
void or_cm31_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]);
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]);

/* ---- circle (src/circle.rs) ----
 * A coset is passed as uint32_t[5] = {init.x, init.y, step.x, step.y, log_size}. */
enum { OR_SUBGROUP = 0, OR_HALF_COSET = 1, OR_HALF_ODDS = 2 };
void or_coset(int kind, uint32_t log_size, uint32_t coset[5]);
void or_eval_domain(uint32_t log_size, uint32_t coset[5]);
void or_coset_at(const uint32_t coset[5], uint64_t i, uint32_t out[2]);
void or_circle_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]);

/* ---- blake2s (src/channel.rs:216-298; standard RFC 7693 Blake2s-256) ---- */
void or_blake2s(const uint8_t *in, size_t len, uint8_t out[32]);
void or_blake2s_single(const uint8_t *in, size_t len, uint8_t out[32]);

/* ---- twiddles (src/circle.rs:284-347, cuda/twiddles.cu:80-142) ---- */
void or_compute_twiddles(const uint32_t coset[5], uint32_t *line_tw,
                         uint32_t *circle_tw, uint32_t *offsets, uint32_t *sizes);
void or_batch_inverse(const uint32_t *in, uint32_t *out, size_t n);
void or_fold_twiddles(const uint32_t coset[5], int extract_y, uint32_t *out);

/* ---- NTT (cuda/circle_ntt.cu:74-117,157-165,248-312) ---- */
void or_ntt_layer(uint32_t *data, const uint32_t *tw, uint32_t layer_idx,
                  uint64_t n, int forward);
void or_ntt_evaluate(uint32_t *data, const uint32_t *line_tw,
                     const uint32_t *circle_tw, const uint32_t *offsets,
                     uint32_t n_line_layers, uint64_t n);
void or_ntt_interpolate(uint32_t *data, const uint32_t *line_itw,
                        const uint32_t *circle_itw, const uint32_t *offsets,
                        uint32_t n_line_layers, uint64_t n);
void or_bit_reverse_m31(uint32_t *data, uint32_t log_n);

/* ---- AIR + constraint (src/air.rs:16-45, cuda/constraint.cu:10-36,81-90) ---- */
void or_fibonacci_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t *out);
void or_zero_pad(const uint32_t *src, uint32_t *dst, uint64_t src_n, uint64_t dst_n);
void or_fibonacci_quotient(const uint32_t *trace, uint32_t *o0, uint32_t *o1,
                           uint32_t *o2, uint32_t *o3, const uint32_t alpha[4],
                           uint64_t n);

/* ---- FRI (cuda/fri.cu:58-126) ---- */
void or_fold_line(const uint32_t *const in[4], const uint32_t *tw,
                  uint32_t *const out[4], const uint32_t alpha[4], uint64_t half_n);
void or_fold_circle_into_line(uint32_t *const dst[4], const uint32_t *const src[4],
                              const uint32_t *tw, const uint32_t alpha[4],
                              const uint32_t alpha_sq[4], uint64_t half_n);

/* ---- Merkle (cuda/blake2s.cu:13-87, src/merkle.rs:520-593,841-884) ---- */
void or_merkle_leaf_hashes(const uint32_t *const *cols, uint32_t n_cols,
                           uint64_t n_leaves, uint32_t *hashes);
void or_merkle_hash_nodes(const uint32_t *children, uint32_t *parents,
                          uint64_t n_parents);
void or_merkle_root(const uint32_t *const *cols, uint32_t n_cols,
                    uint32_t log_n_leaves, uint32_t root[8]);
int or_verify_auth_path(const uint32_t root[8], const uint32_t leaf_hash[8],
                        uint64_t index, const uint32_t *path, uint32_t depth);

/* ---- channel (src/channel.rs:15-185) ---- */
typedef struct {
    uint8_t state[32];
    uint32_t n_draws;
} or_channel;
void or_channel_init(or_channel *ch);
void or_channel_mix_digest(or_channel *ch, const uint32_t digest[8]);
void or_channel_mix_felts(or_channel *ch, const uint32_t *felts, size_t n_felts);
void or_channel_draw_felt(or_channel *ch, uint32_t out[4]);
uint64_t or_channel_draw_number(or_channel *ch, uint64_t bound);
void or_channel_mix_u64(or_channel *ch, uint64_t value);
int or_channel_verify_pow_nonce(const or_channel *ch, uint32_t pow_bits, uint64_t nonce);
uint64_t or_grind_pow(const or_channel *ch, uint32_t pow_bits);
/* stwo plugin column ops: cuda/accumulate_lift.cu:14-40, accumulation_ops.rs:44-85 (one channel) */
void or_accumulate_lift(uint32_t *col, const uint32_t *curr, uint64_t col_n, uint32_t log_ratio);
void or_lift_and_accumulate(uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols);
/* the stwo plugin's nonce family (hi << 32) | low, low < 2^20 (grind_ops.rs:71-128) */
uint64_t or_grind_pow_stwo(const or_channel *ch, uint32_t pow_bits);
/* the same search over a caller-given prefixed digest; m31_out reduces the
 * hash words mod P first (Blake2sChannelGeneric<true>, grind_ops.rs:154-245) */
uint32_t or_pow_tz_words(const uint32_t prefix[8], uint64_t nonce, int m31_out);
uint64_t or_grind_pow_stwo_words(const uint32_t prefix[8], uint32_t pow_bits, int m31_out);

/* ---- prove / verify (src/prover.rs:264-619, src/verifier.rs:39-310,
 *      src/bin/stark_cli.rs:1051-1110) ---- */
int or_prove(uint32_t a, uint32_t b, uint32_t log_n, uint32_t blowup_bits,
             uint32_t n_queries, uint32_t pow_bits, uint8_t **out, size_t *out_len);
int or_verify(const uint8_t *bytes, size_t len, uint32_t blowup_bits,
              uint32_t n_queries, char *err, size_t err_cap);
size_t or_proof_size(uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries,
                     uint32_t pow_bits);
void or_free(void *p);

/* ---- synthetic degree-2 multi-column AIR (build-defined; no reference
 *      counterpart): W columns, K constraints C_k = c_k t_a t_b' + d_k t_e,
 *      proof bytes KRAK v3 (definition above or_d2_value in vx_oracle.c) ---- */
uint32_t or_d2_value(uint64_t seed, uint32_t col, uint64_t row);
void or_d2_trace(uint64_t seed, uint32_t n_cols, uint32_t log_n, uint32_t *out);
void or_d2_column(uint64_t seed, uint32_t col, uint32_t log_n, uint32_t *out);
void or_d2_coeffs(uint64_t air_seed, uint32_t n_cons, uint32_t *cd);
void or_d2_quotient(const uint32_t *eval, uint32_t n_cols, uint32_t le, uint32_t n_cons, uint64_t air_seed,
                    const uint32_t alpha[4], uint32_t *q0, uint32_t *q1, uint32_t *q2, uint32_t *q3);
int or_d2_prove(uint64_t seed, uint32_t log_n, uint32_t n_cols, uint32_t n_cons, uint64_t air_seed,
                uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits, uint8_t **out, size_t *out_len);
int or_d2_verify(const uint8_t *bytes, size_t len, uint32_t blowup_bits, uint32_t n_queries, char *err,
                 size_t err_cap);
size_t or_d2_proof_size(uint32_t log_n, uint32_t n_cols, uint32_t blowup_bits, uint32_t n_queries);

/* ---- batched uneven columns: per-column LDE (src/prover.rs:280-318) and the
 *      lifted Merkle tree (cuda/merkle_leaves_lifted.cu:26-98) ---- */
void or_lde(const uint32_t *in, uint32_t log_n, uint32_t blowup_bits, uint32_t *out);
/* Outputs idx[0..n_idx) of or_ntt_evaluate on the domain of coset dom applied
 * to the zero-padded coefficients c[0 .. 2^log_n), O(2^log_n) per group of
 * outputs sharing bits >= 2 (closed-form twiddles, product basis). */
/* Cairo VM AIR quotient (cuda/cairo_constraint.cu): 34 trace columns,
 * interaction columns as QM31 SoA, 40 (full / slab) or 35 (chunk) weighted
 * constraints, divided by Z_H. */
void or_cairo_vanishing(const uint32_t init[2], const uint32_t step[2], uint32_t *out, uint32_t log_eval,
                        uint32_t log_n, int inverse);
void or_axpy_m31(uint32_t s, const uint32_t *x, uint32_t *y, uint64_t n);
void or_cairo_quotient(const uint32_t *const cols[34], const uint32_t *const inter[32], uint32_t *const out[4],
                       const uint32_t alpha[160], const uint32_t *vh_inv, const uint32_t *trans,
                       const uint32_t chal[24], uint64_t n, uint32_t blowup_step, uint64_t chunk_offset,
                       int slab_mode);
void or_cairo_quotient_chunk(const uint32_t *const cols[34], const uint32_t *const inter[12],
                             uint32_t *const out[4], const uint32_t alpha[140], const uint32_t *vh_inv,
                             const uint32_t *trans, const uint32_t chal[24], uint64_t offset, uint64_t chunk_n,
                             uint64_t global_n, uint32_t blowup_step);
void or_ntt_evaluate_at(const uint32_t *c, uint32_t log_n, const uint32_t dom5[5], const uint64_t *idx,
                        uint32_t n_idx, uint32_t *out);
void or_build_leaves_lifted(const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                            uint32_t lifting_log, uint32_t *hashes);
void or_merkle_root_lifted(const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                           uint32_t root[8]);

/* ---- batched-columns proof (KRAK v4): per-column bit-reversed LDE, lifted
 *      tree, random-linear-combination quotient, the Fibonacci FRI ---- */
int or_batch_prove(uint64_t seed, const uint32_t *log_sizes, uint32_t n_cols, uint32_t blowup_bits,
                   uint32_t n_queries, uint32_t pow_bits, uint8_t **out, size_t *out_len);
int or_batch_verify(const uint8_t *bytes, size_t len, uint32_t blowup_bits, uint32_t n_queries, char *err,
                    size_t err_cap);
size_t or_batch_proof_size(const uint32_t *log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t n_queries);

/* ---- stwo-basis circle NTT (stwo plugin PolyOps): src/ntt.rs:296-354,
 *      cuda/circle_ntt_stwo.cu, src/oods.rs:133-161; log_n >= 2 ---- */
void or_stwo_twiddles(uint32_t log_n, uint32_t *fwd, uint32_t *inv);
void or_stwo_ntt_evaluate(uint32_t *data, const uint32_t *tw, uint32_t log_n);
void or_stwo_ntt_interpolate(uint32_t *data, const uint32_t *itw, uint32_t log_n);
/* half-coset natural <-> canonic bit-reversed order (src/circle.rs:192-220) */
void or_permute_hc_to_canonic_brt(const uint32_t *src, uint32_t *dst, uint32_t log_n);
void or_permute_canonic_brt_to_hc_natural(const uint32_t *src, uint32_t *dst, uint32_t log_n);
/* the OODS step of src/oods.rs: draw z, z_next = z * step(trace_log),
 * evaluate every column at both, mix the values (all at z first) */
void or_oods_sample(or_channel *ch, const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                    uint32_t trace_log, uint32_t points[16], uint32_t *values);
void or_stwo_eval_at_point(const uint32_t *coeffs, uint32_t log_n, const uint32_t zx[4], const uint32_t zy[4],
                           uint32_t out[4]);
/* cuda/circle_ntt.cu:186-231,429-468 (cuda_eval_at_point, folding factors given);
 * cuda/barycentric_eval.cu:41-110 (per-block partial sums) */
void or_eval_at_point_factors(const uint32_t *coeffs, const uint32_t *factors, uint32_t log_n, uint32_t out[4]);
void or_barycentric_partials(const uint32_t *evals, const uint32_t *weights, uint64_t n, uint32_t n_blocks,
                             uint32_t *out);

/* ---- stwo plugin QuotientOps: cuda/quotient.cu (pointer arrays on the host) ---- */
void or_accumulate_numerators(const uint32_t *const *cols, const uint32_t *col_indices, const uint32_t *b,
                              const uint32_t *c, uint32_t n_batch_cols, uint64_t n_rows, uint32_t *o0,
                              uint32_t *o1, uint32_t *o2, uint32_t *o3);
void or_quotients_combine(const uint32_t *sp_x, const uint32_t *sp_y, const uint32_t *first_linear_acc,
                          const uint32_t *const *numer0, const uint32_t *const *numer1,
                          const uint32_t *const *numer2, const uint32_t *const *numer3,
                          const uint32_t *acc_log_sizes, uint32_t n_accs, const uint32_t *domain_xs,
                          const uint32_t *domain_ys, uint32_t lifting_log_size, uint64_t n_rows, uint32_t *o0,
                          uint32_t *o1, uint32_t *o2, uint32_t *o3);

/* ---- generic constraint evaluation: cuda/constraint_eval.cu:97-407 ----
 * The reference's register bytecode run row by row with a 65,536-entry QM31
 * register file; acc_c[row] += (denom_inv[row >> trace_log] * sum_k
 * coeff_k * constraint_k)_c. Returns 0, or -1 (accumulators unchanged)
 * for a column index out of range, from which the reference kernel's rows
 * return without accumulating, and for an unknown opcode or a truncated
 * instruction, which the reference kernel skips or reads past. */
uint32_t or_bc_offset_row(uint32_t i, uint32_t domain_log, uint32_t eval_log, int32_t offset);
int or_bytecode_constraint_eval(const uint32_t *bytecode, uint32_t n_words, const uint32_t *const *cols,
                                uint32_t n_cols, uint32_t n_rows, uint32_t trace_n_rows, const uint32_t *coeff,
                                const uint32_t *denom_inv, uint32_t *acc0, uint32_t *acc1, uint32_t *acc2,
                                uint32_t *acc3);

/* ---- GKR sum-check operations of the stwo plugin: cuda/gkr.cu (secure
 * columns 4 words per element) ---- */
void or_gkr_fix_first_variable_base(const uint32_t *in, uint32_t *out, const uint32_t r[4], uint64_t n);
void or_gkr_fix_first_variable_secure(const uint32_t *in, uint32_t *out, const uint32_t r[4], uint64_t n);
void or_gkr_gen_eq_evals_pass(uint32_t *buf, const uint32_t y[4], uint64_t cur);
void or_gkr_gen_eq_evals(const uint32_t *y, uint32_t k, const uint32_t v[4], uint32_t *out);
void or_gkr_next_layer(int kind, const uint32_t *in_num, const uint32_t *in_den, uint32_t *out_num,
                       uint32_t *out_den, uint64_t n);
void or_gkr_sum_poly(int kind, const uint32_t *eq, const uint32_t *in_num, const uint32_t *in_den,
                     const uint32_t lambda[4], uint64_t n_terms, uint32_t out[8]);

#ifdef __cplusplus
}
#endif
#endif
