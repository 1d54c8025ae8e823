/*
 * vortex_hip.h -- C-ABI of the MI355X (gfx950) Circle-STARK prover hot path.
 *
 * Drop-in replacement for the CUDA FFI of garrick247/VortexSTARK
 * (src/cuda/ffi.rs, resolved against libvortex_cuda.a built from the cuda/ kernels).
 * Each entry point names the reference declaration it replaces. Conventions:
 *   - same argument order and meaning as the reference `cuda_*` function;
 *   - the `cuda_` prefix becomes `vx_`; a trailing hipStream_t selects the
 *     stream (NULL = the legacy null stream);
 *   - every function returns int: 0 = success, otherwise a vx_status code,
 *     with a message from vx_last_error(); the reference returns void and
 *     panics on the next failing runtime call;
 *   - the caller owns and allocates all device buffers, outputs included;
 *   - scalars given as `const uint32_t *alpha` ([4] QM31 words) are read from
 *     HOST memory at call time, as in the reference (cuda/fri.cu:40-52);
 *   - sizes are element counts (uint32_t in the reference; uint64_t is used
 *     where 2^31-element domains need it).
 * Arithmetic is exact over GF(2^31 - 1): results are bit-identical to the
 * reference kernels for the same inputs.
 */
#ifndef VORTEX_HIP_H
#define VORTEX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *vx_stream_t; /* == hipStream_t */

enum vx_status {
    VX_OK = 0,
    VX_ERR_ARG = 1,
    VX_ERR_HIP = 2,
    VX_ERR_ALLOC = 3,
    VX_ERR_VERIFY = 4,
    VX_ERR_INTERNAL = 5,
};

/* Message for the last non-zero status returned on this thread. */
const char *vx_last_error(void);
/* Replaces cuda_device_sync (cuda/fields.cu:44-46; ffi.rs). */
int vx_device_sync(void);

/* Runtime helpers replacing the CUDA runtime bindings the host side uses
 * (src/cuda/ffi.rs:6-28: cudaMalloc, cudaFree, cudaMemcpy[Async],
 * cudaMemset, cudaStreamCreate/Synchronize/Destroy), so callers share this
 * library's HIP runtime. kind: 1 = host->device, 2 = device->host,
 * 3 = device->device (cudaMemcpyKind / hipMemcpyKind values). */
int vx_malloc(void **d_ptr, size_t bytes);
int vx_free(void *d_ptr);
int vx_memcpy(void *dst, const void *src, size_t bytes, int kind);
int vx_memcpy_async(void *dst, const void *src, size_t bytes, int kind, vx_stream_t stream);
int vx_memset(void *d_ptr, int value, size_t bytes);
int vx_stream_create(vx_stream_t *out);
int vx_stream_synchronize(vx_stream_t stream);
int vx_stream_destroy(vx_stream_t stream);

/* ------------------------------------------------------------------ NTT */

/* Replaces cuda_circle_ntt_evaluate (src/cuda/ffi.rs:261-269,
 * cuda/circle_ntt.cu:248-275). Twiddle tables in the reference layout
 * (src/circle.rs:284-347); h_layer_offsets/h_layer_sizes are HOST arrays.
 * Like the reference, returns after the stream has finished. */
int vx_circle_ntt_evaluate(uint32_t *d_data, const uint32_t *d_twiddles, const uint32_t *d_circle_twids,
                           const uint32_t *h_layer_offsets, const uint32_t *h_layer_sizes,
                           uint32_t n_line_layers, uint32_t n, vx_stream_t stream);

/* Replaces cuda_circle_ntt_interpolate (ffi.rs:271-279, circle_ntt.cu:277-312),
 * including the final 2^((30 log n) mod 31) scale. Synchronous like the reference. */
int vx_circle_ntt_interpolate(uint32_t *d_data, const uint32_t *d_itwiddles, const uint32_t *d_circle_itwids,
                              const uint32_t *h_layer_offsets, const uint32_t *h_layer_sizes,
                              uint32_t n_line_layers, uint32_t n, vx_stream_t stream);

/* Replaces cuda_circle_ntt_evaluate_batch / _interpolate_batch
 * (ffi.rs:281-301, callers src/ntt.rs:226-240,270-281): d_columns is a
 * DEVICE array of n_cols device column pointers, as the reference uploads it
 * (a host array is accepted too: the pointers are read back with the copy
 * direction inferred from the address). Columns of equal size n, 16-byte
 * aligned; each pass is one launch over all columns. Synchronous. */
int vx_circle_ntt_evaluate_batch(uint32_t *const *d_columns, const uint32_t *d_twiddles,
                                 const uint32_t *d_circle_twids, const uint32_t *h_layer_offsets,
                                 const uint32_t *h_layer_sizes, uint32_t n_line_layers, uint32_t n,
                                 uint32_t n_cols, vx_stream_t stream);
int vx_circle_ntt_interpolate_batch(uint32_t *const *d_columns, const uint32_t *d_itwiddles,
                                    const uint32_t *d_circle_itwids, const uint32_t *h_layer_offsets,
                                    const uint32_t *h_layer_sizes, uint32_t n_line_layers, uint32_t n,
                                    uint32_t n_cols, vx_stream_t stream);
/* Replace cuda_circle_ntt_evaluate_batch_forge / _interpolate_batch_forge
 * (ffi.rs:407-428, the reference's default-feature build, src/ntt.rs:
 * 208-222): the same transforms over a HOST array of device column pointers. */
int vx_circle_ntt_evaluate_batch_host(uint32_t *const *h_columns, const uint32_t *d_twiddles,
                                      const uint32_t *d_circle_twids, const uint32_t *h_layer_offsets,
                                      const uint32_t *h_layer_sizes, uint32_t n_line_layers, uint32_t n,
                                      uint32_t n_cols, vx_stream_t stream);
int vx_circle_ntt_interpolate_batch_host(uint32_t *const *h_columns, const uint32_t *d_itwiddles,
                                         const uint32_t *d_circle_itwids, const uint32_t *h_layer_offsets,
                                         const uint32_t *h_layer_sizes, uint32_t n_line_layers, uint32_t n,
                                         uint32_t n_cols, vx_stream_t stream);

/* Replaces cuda_circle_ntt_layer (ffi.rs:303-309, circle_ntt.cu:405-417):
 * one butterfly layer with stride 2^layer_idx (0 = circle layer). Async. */
int vx_circle_ntt_layer(uint32_t *d_data, const uint32_t *d_twiddles, uint32_t layer_idx, uint32_t n,
                        int forward, vx_stream_t stream);

/* Fused zero_pad + evaluate (src/prover.rs:303-318 = cuda_zero_pad then
 * cuda_circle_ntt_evaluate): d_eval[0..n) = evaluate(coeffs zero-padded to n)
 * on the domain whose tables are given. The first log2(n / coeff_n) layers
 * only replicate the coefficients and are skipped. Async. */
int vx_circle_ntt_evaluate_lde(uint32_t *d_eval, const uint32_t *d_coeffs, uint32_t coeff_n,
                               const uint32_t *d_twiddles, const uint32_t *d_circle_twids,
                               const uint32_t *h_layer_offsets, uint32_t n_line_layers, uint32_t n,
                               vx_stream_t stream);

/* Replaces cuda_bit_reverse_m31 (ffi.rs:311, circle_ntt.cu:234-244,422-427). Async. */
int vx_bit_reverse_m31(uint32_t *d_data, uint32_t log_n, vx_stream_t stream);
/* Replaces cuda_bit_reverse_qm31 (ffi.rs:1184, cuda/bit_reverse_wide.cu:21-52):
 * out-of-place, 4 words per element. Async. */
int vx_bit_reverse_qm31(const uint32_t *d_in, uint32_t *d_out, uint32_t n, uint32_t log_n, vx_stream_t stream);

/* --------------------------------------------- stwo-basis NTT (stwo plugin)
 * The stwo Backend plugin's PolyOps transforms (crates/cuda-backend/src/
 * poly_ops.rs:361-381 and its evaluate/extend paths), in stwo's basis
 * f = sum c_i * prod_{bit b of i} m_b (m_0 = y, m_1 = x, m_b = 2 m_{b-1}^2 - 1),
 * values in bit-reversed CircleDomain order. Twiddles are stwo's flat buffer
 * of n/2 words (src/ntt.rs:296-354): line layer k at n/2 - 2^(log2(n)-1-k),
 * then a pad word 1; the circle twiddles are derived from line layer 0.
 * n is a power of two in [4, 2^30]; columns 16-byte aligned. At n = 4 the
 * circle layer uses the pad word as its y twiddle, as the reference does. */

/* stwo forward (d_tw) and, if d_itw is not NULL, inverse twiddle buffers of
 * 2^(log_n - 1) words each for a 2^log_n-point domain, computed on the device
 * (replaces the host loop of StwoNttCache::new, src/ntt.rs:296-354, and the
 * plugin's CpuBackend::precompute_twiddles upload, poly_ops.rs:102-127). Async. */
int vx_stwo_twiddles(uint32_t log_n, uint32_t *d_tw, uint32_t *d_itw, vx_stream_t stream);

/* Replaces cuda_stwo_ntt_evaluate (src/cuda/ffi.rs:235-236,
 * cuda/circle_ntt_stwo.cu:268-337): coefficients -> values. Synchronous like
 * the reference. */
int vx_stwo_ntt_evaluate(uint32_t *d_data, const uint32_t *d_twiddles, uint32_t n, vx_stream_t stream);
/* Replaces cuda_stwo_ntt_interpolate (ffi.rs:237-238, circle_ntt_stwo.cu:
 * 339-402): values -> coefficients, including the 2^((30 log n) mod 31) scale.
 * Synchronous. */
int vx_stwo_ntt_interpolate(uint32_t *d_data, const uint32_t *d_itwiddles, uint32_t n, vx_stream_t stream);
/* Replace cuda_stwo_ntt_batch_evaluate / _batch_interpolate (ffi.rs:239-256,
 * cuda/circle_ntt_stwo_batch.cu:239-372): n_cols columns of n words at the
 * device addresses in the HOST array h_columns, each pass one launch for all
 * columns. n_cols = 0 is a no-op. Synchronous. */
int vx_stwo_ntt_batch_evaluate(uint32_t *const *h_columns, const uint32_t *d_twiddles, uint32_t n, uint32_t n_cols,
                               vx_stream_t stream);
int vx_stwo_ntt_batch_interpolate(uint32_t *const *h_columns, const uint32_t *d_itwiddles, uint32_t n,
                                  uint32_t n_cols, vx_stream_t stream);
/* The plugin's PolyOps::evaluate on a larger domain (poly_ops.rs:652-672:
 * extend = zero-fill + copy of the coefficients, then cuda_stwo_ntt_evaluate
 * over all n words) in one call: d_eval[0..n) = evaluate(d_coeffs[0..coeff_n)
 * zero-padded to n), d_twiddles being the n-point domain's. The top
 * log2(n / coeff_n) layers only replicate the coefficients and are not run,
 * and there is no zero-fill or copy pass. coeff_n a power of two in [2, n];
 * d_eval must not alias d_coeffs. Async. */
int vx_stwo_ntt_evaluate_lde(uint32_t *d_eval, const uint32_t *d_coeffs, uint32_t coeff_n,
                             const uint32_t *d_twiddles, uint32_t n, vx_stream_t stream);
/* The same for the plugin's evaluate_polynomials groups (poly_ops.rs:674-760:
 * per target size, allocate + zero-fill + copy each column, then
 * cuda_stwo_ntt_batch_evaluate): n_cols columns of coeff_n coefficients
 * (HOST array h_coeffs of device pointers) into n-word evaluations (HOST
 * array h_eval), every pass one launch for all columns. Synchronous. */
int vx_stwo_ntt_evaluate_lde_batch(uint32_t *const *h_eval, const uint32_t *const *h_coeffs, uint32_t coeff_n,
                                   uint32_t n_cols, const uint32_t *d_twiddles, uint32_t n, vx_stream_t stream);
/* No counterpart in the reference: the same transforms, asynchronous, over
 * n_cols (1..65535) columns `stride` words apart (stride >= n when n_cols > 1),
 * for callers that keep columns in one allocation and order work by stream. */
int vx_stwo_ntt_evaluate_async(uint32_t *d_data, const uint32_t *d_twiddles, uint32_t n, uint32_t n_cols,
                               uint64_t stride, vx_stream_t stream);
int vx_stwo_ntt_interpolate_async(uint32_t *d_data, const uint32_t *d_itwiddles, uint32_t n, uint32_t n_cols,
                                  uint64_t stride, vx_stream_t stream);

/* Replaces cuda_eval_at_point (src/cuda/ffi.rs:334-341, cuda/circle_ntt.cu:
 * 186-231,429-468; the plugin's PolyOps::eval_at_point, crates/cuda-backend/
 * src/poly_ops.rs:382-443): f(z) of n stwo-basis coefficients, given the
 * log2(n) QM31 folding factors (device, 4 words each, factors[0] folding the
 * top coefficient bit: the mappings [z.y, z.x, 2 z.x^2 - 1, ...] reversed).
 * The QM31 result is written to the HOST array h_result[4]. Each scratch
 * buffer needs 4 words for n < 2^14, else max(80, n / 1024) words (the
 * reference's callers pass 2n). n is a power of two; coefficients 16-byte
 * aligned. Synchronous. */
int vx_eval_at_point(const uint32_t *d_coeffs, const uint32_t *d_folding_factors, uint32_t *h_result, uint32_t n,
                     uint32_t *d_scratch1, uint32_t *d_scratch2, vx_stream_t stream);
/* Many polynomials at one point (the OODS step). Replaces
 * cuda_batch_eval_at_point (cuda/batch_eval_at_point.cu:11-100), which is
 * limited to 64 coefficients per polynomial and marks larger ones with
 * 0xFFFFFFFF; here any power-of-two size up to 2^max_log_n. Column k has
 * 2^h_log_sizes[k] coefficients and uses the last h_log_sizes[k] of the
 * max_log_n factors (d_folding_factors, the factors of a 2^max_log_n
 * polynomial, as for vx_eval_at_point). h_cols and h_log_sizes are HOST
 * arrays (the reference passes device arrays); columns of 4 or more words
 * 16-byte aligned. d_results: n_cols QM31 (device). One launch for every
 * 2^14-coefficient chunk of every column, then one per fold level.
 * Synchronous. */
int vx_eval_at_point_batch(const uint32_t *const *h_cols, const uint32_t *h_log_sizes, uint32_t n_cols,
                           const uint32_t *d_folding_factors, uint32_t max_log_n, uint32_t *d_results,
                           vx_stream_t stream);
/* The OODS step (src/oods.rs:1-19): on a DEVICE channel of 9 words (the
 * 32-byte state as 8 LE words, then n_draws), draw t (Channel::draw_felt)
 * and the point z = ((1 - t^2)/(1 + t^2), 2t/(1 + t^2)) (OodsPoint::
 * from_channel, :44-52), z_next = z * step with step the generator of the
 * 2^trace_log_n subgroup (next_step, :74-81); evaluate every stwo-basis
 * coefficient column (HOST arrays of device pointers and log sizes <= 30) at
 * z and at z_next; mix the 2 n_cols values into the channel
 * (flatten_for_channel order, :395-406: all f_k(z), then all f_k(z_next);
 * Channel::mix_felts, src/channel.rs:46-59). Writes d_points[16] =
 * (z.x, z.y, z_next.x, z_next.y) and d_values[8 n_cols] in that order; the
 * channel is updated in place. No counterpart kernel in the reference (its
 * OODS step runs on the host); the per-column evaluation is
 * vx_eval_at_point_batch. Synchronous. */
int vx_oods_sample(uint32_t *d_channel, const uint32_t *const *h_cols, const uint32_t *h_log_sizes, uint32_t n_cols,
                   uint32_t trace_log_n, uint32_t *d_points, uint32_t *d_values, vx_stream_t stream);
/* Replaces cuda_barycentric_eval (ffi.rs:1037-1043, cuda/barycentric_eval.cu:
 * 41-110): d_out[4b..4b+4) = sum of d_evals[i] * d_weights[i] (QM31, 4 words
 * each) over i = b*256 + t + k*n_blocks*256 (t < 256), the reference's
 * partition, so each partial equals the reference's; the caller adds the
 * n_blocks partials. Weights 16-byte aligned. Synchronous. */
int vx_barycentric_eval(const uint32_t *d_evals, const uint32_t *d_weights, uint32_t n, uint32_t *d_out,
                        uint32_t n_blocks, vx_stream_t stream);

/* ------------------------------------- stwo plugin QuotientOps (quotients)
 * Replaces cuda_accumulate_numerators (src/cuda/ffi.rs:1059-1070,
 * cuda/quotient.cu:13-50,240-257), called by the plugin's
 * QuotientOps::accumulate_numerators (crates/cuda-backend/src/
 * quotient_ops.rs:120-208): for row < n_rows,
 *   out[row] = sum_i (c_i * d_col_ptrs[d_col_indices[i]][row] - b_i)
 * with b_i, c_i QM31 (4 words each) and the result as 4 SoA columns.
 * d_col_ptrs, d_col_indices, d_b, d_c are DEVICE arrays. Async. */
int vx_accumulate_numerators(const uint32_t *const *d_col_ptrs, const uint32_t *d_col_indices, const uint32_t *d_b,
                             const uint32_t *d_c, uint32_t n_batch_cols, uint32_t n_rows, uint32_t *out0,
                             uint32_t *out1, uint32_t *out2, uint32_t *out3, vx_stream_t stream);
/* Replaces cuda_accumulate_numerators_dual (ffi.rs:1072-1089, quotient.cu:
 * 56-131,259-282): the z sums over all n_cols_z columns and, from the same
 * column reads, the z_next sums over the columns whose d_zn_coeff_idx[i] is
 * not 0xFFFFFFFF, with coefficients b_zn / c_zn at that index. The reference
 * stages at most 128 columns in shared memory; here any count. Async. */
int vx_accumulate_numerators_dual(const uint32_t *const *d_col_ptrs, const uint32_t *d_col_indices_z,
                                  const uint32_t *d_zn_coeff_idx, const uint32_t *d_b_z, const uint32_t *d_c_z,
                                  const uint32_t *d_b_zn, const uint32_t *d_c_zn, uint32_t n_cols_z, uint32_t n_rows,
                                  uint32_t *out_z0, uint32_t *out_z1, uint32_t *out_z2, uint32_t *out_z3,
                                  uint32_t *out_zn0, uint32_t *out_zn1, uint32_t *out_zn2, uint32_t *out_zn3,
                                  vx_stream_t stream);
/* Replaces cuda_compute_quotients_combine (ffi.rs:1091-1124, quotient.cu:
 * 153-233,284-312; QuotientOps::compute_quotients_and_combine,
 * quotient_ops.rs:210-342): for row < n_rows with domain point
 * d = (d_xs[row], d_ys[row]),
 *   out[row] = sum_j (numer_j[lifted(row)] - first_linear_acc_j * d.y) / den_j,
 *   den_j = (Re p_j.x - d.x) Im p_j.y - (Re p_j.y - d.y) Im p_j.x   (CM31),
 *   lifted(row) = ((row >> (lifting_log_size - acc_log_sizes[j] + 1)) << 1) + (row & 1),
 * each CM31 half of the QM31 numerator divided by den_j (zero inverts to
 * zero). Sample points, accumulators, pointer arrays and log sizes are
 * DEVICE arrays; n_rows <= 2^lifting_log_size. lifted(row) is clamped to
 * the accumulator's last row, so an accumulator of log size 0, or one larger
 * than the lifting domain, is read in bounds. Async. */
int vx_compute_quotients_combine(const uint32_t *d_sample_points_x, const uint32_t *d_sample_points_y,
                                 const uint32_t *d_first_linear_acc, const uint32_t *const *d_numer_ptrs0,
                                 const uint32_t *const *d_numer_ptrs1, const uint32_t *const *d_numer_ptrs2,
                                 const uint32_t *const *d_numer_ptrs3, const uint32_t *d_acc_log_sizes,
                                 uint32_t n_accs, const uint32_t *d_domain_xs, const uint32_t *d_domain_ys,
                                 uint32_t lifting_log_size, uint32_t n_rows, uint32_t *out0, uint32_t *out1,
                                 uint32_t *out2, uint32_t *out3, vx_stream_t stream);

/* ------------------------------------- stwo plugin column ops (columns)
 * Replace cuda_m31_add / cuda_m31_mul (src/cuda/ffi.rs:212-213; the plugin's
 * AccumulationOps::accumulate, accumulation_ops.rs:11-31): out = a + b (or
 * a * b) element-wise over n M31 words. Async. */
int vx_m31_add(const uint32_t *d_a, const uint32_t *d_b, uint32_t *d_out, uint32_t n, vx_stream_t stream);
int vx_m31_mul(const uint32_t *d_a, const uint32_t *d_b, uint32_t *d_out, uint32_t n, vx_stream_t stream);
/* Replaces cuda_accumulate_lift (ffi.rs:1186, cuda/accumulate_lift.cu:14-40):
 * col[i] += curr[((i >> (log_ratio + 1)) << 1) | (i & 1)] for i < col_n. Async. */
int vx_accumulate_lift(uint32_t *d_col, const uint32_t *d_curr, uint32_t col_n, uint32_t log_ratio,
                       vx_stream_t stream);
/* No counterpart as one call in the reference: the plugin's
 * AccumulationOps::lift_and_accumulate (accumulation_ops.rs:44-85), which
 * runs accumulate_lift per column and channel with a synchronisation after
 * each column. Columns in ascending log size (h_log_sizes, 1..31, HOST);
 * h_cols is a HOST array of 4 x n_cols device pointers, channel-major
 * (h_cols[c * n_cols + k] = channel c of column k). The largest (last)
 * column receives the lifted sum of all columns in one pass; the others
 * are only read. Synchronous. */
int vx_lift_and_accumulate(uint32_t *const *h_cols, const uint32_t *h_log_sizes, uint32_t n_cols,
                           vx_stream_t stream);
/* Replace cuda_gather_u32 / cuda_gather_u256 (ffi.rs:980-984, cuda/gather.cu)
 * and their bounds-checked forms cuda_gather_u32_forge / _u256_forge
 * (ffi.rs:986-1002, cuda/forge/gather_forge.cu:82-116): dst[i] = src[idx[i]]
 * with 1 or 8 words per element; the checked forms skip elements whose
 * source words are not all below src_len (words). Async. */
int vx_gather_u32(const uint32_t *d_src, const uint32_t *d_idx, uint32_t *d_dst, uint32_t n, vx_stream_t stream);
int vx_gather_u32_checked(const uint32_t *d_src, const uint32_t *d_idx, uint32_t *d_dst, uint32_t n,
                          uint32_t src_len, vx_stream_t stream);
int vx_gather_u256(const uint32_t *d_src, const uint32_t *d_idx, uint32_t *d_dst, uint32_t n, vx_stream_t stream);
int vx_gather_u256_checked(const uint32_t *d_src, const uint32_t *d_idx, uint32_t *d_dst, uint32_t n,
                           uint32_t src_len, vx_stream_t stream);

/* Replace cuda_permute_hc_to_canonic_brt and
 * cuda_permute_canonic_brt_to_hc_natural (ffi.rs:317-330, cuda/permute.cu:
 * 30-57; Coset::permute_hc_natural_to_canonic_brt and its inverse,
 * src/circle.rs:192-220): for k < n = 2^log_n, with
 * cn = k / 2 (k even) or n - 1 - k / 2 (k odd) and j = bit-reverse(cn),
 * the first sets dst[j] = src[k], the second dst[k] = src[j]. Out of place
 * (src != dst). log_n <= 31. Async. */
int vx_permute_hc_to_canonic_brt(const uint32_t *d_src, uint32_t *d_dst, uint32_t n, uint32_t log_n,
                                 vx_stream_t stream);
int vx_permute_canonic_brt_to_hc_natural(const uint32_t *d_src, uint32_t *d_dst, uint32_t n, uint32_t log_n,
                                         vx_stream_t stream);
/* Replaces cuda_pack_leaves (ffi.rs:1187-1195, cuda/pack_leaves.cu:13-35; the
 * plugin's PackLeavesOps::pack_leaves_input, merkle_ops.rs:197-228): 4 M31
 * columns of N words (N a multiple of 16) into 64 columns of N / 16 words,
 * out[16 c + (p mod 16)][p / 16] = in[c][p]. d_input_col_ptrs (4) and
 * d_output_col_ptrs (64) are DEVICE arrays of device pointers. Async. */
int vx_pack_leaves(const uint32_t *const *d_input_col_ptrs, uint32_t *const *d_output_col_ptrs, uint32_t n,
                   vx_stream_t stream);

/* ------------------------------------------------------------- twiddles */

/* Replaces cuda_compute_coset_points (ffi.rs:453-458, cuda/twiddles.cu:80-105,182-194). */
int vx_compute_coset_points(uint32_t initial_x, uint32_t initial_y, uint32_t step_x, uint32_t step_y,
                            uint32_t *d_out_x, uint32_t *d_out_y, uint32_t n, vx_stream_t stream);
/* Replaces cuda_extract_and_squash (ffi.rs:462-467, twiddles.cu:127-142,204-213). */
int vx_extract_and_squash(const uint32_t *d_in, uint32_t *d_twiddle_out, uint32_t *d_squash_out,
                          uint32_t half_n, vx_stream_t stream);
/* Replaces cuda_squash_x (ffi.rs:460, twiddles.cu:109-123,196-202). */
int vx_squash_x(const uint32_t *d_in, uint32_t *d_out, uint32_t out_n, vx_stream_t stream);
/* Replaces cuda_batch_inverse_m31[_stream] (ffi.rs:440,451, cuda/batch_inverse.cu:8-57);
 * inverse(0) = 0. In-place (d_in == d_out) is allowed. */
int vx_batch_inverse_m31(const uint32_t *d_in, uint32_t *d_out, uint32_t n, vx_stream_t stream);
/* Replaces cuda_compute_fold_twiddle_sources[_stream] (ffi.rs:432-449,
 * twiddles.cu:23-52): out[i] = coord(init * step^bitrev_log_n(2i)), i < n/2. */
int vx_compute_fold_twiddle_sources(uint32_t initial_x, uint32_t initial_y, uint32_t step_x, uint32_t step_y,
                                    uint32_t *d_out, uint32_t n, uint32_t log_n, int extract_y,
                                    vx_stream_t stream);

/* ---------------------------------------------------- AIR / constraints */

/* Device counterpart of air::fibonacci_trace_parallel (src/air.rs:136-189):
 * d_out[i] = Fib(a, b)[i] for i < 2^log_n. */
int vx_fibonacci_trace(uint32_t a, uint32_t b, uint32_t *d_out, uint32_t log_n, vx_stream_t stream);
/* Replaces cuda_fibonacci_quotient (ffi.rs:472-477, cuda/constraint.cu:10-36,94-105). */
int vx_fibonacci_quotient(const uint32_t *d_trace, uint32_t *d_out0, uint32_t *d_out1, uint32_t *d_out2,
                          uint32_t *d_out3, const uint32_t *h_alpha, uint32_t n, vx_stream_t stream);
/* Replaces cuda_fibonacci_quotient_chunk[_stream] (ffi.rs:486-503, constraint.cu:38-64,121-153). */
int vx_fibonacci_quotient_chunk(const uint32_t *d_trace, uint32_t *d_out0, uint32_t *d_out1, uint32_t *d_out2,
                                uint32_t *d_out3, const uint32_t *h_alpha, uint32_t offset, uint32_t chunk_n,
                                uint32_t global_n, vx_stream_t stream);
/* Replaces cuda_zero_pad (ffi.rs:479-484, constraint.cu:81-90,108-117). */
int vx_zero_pad(const uint32_t *d_src, uint32_t *d_dst, uint32_t src_n, uint32_t dst_n, vx_stream_t stream);

/* ------------------------------------------------------------------ FRI */

/* Replaces cuda_fold_line_soa (ffi.rs:346-352, cuda/fri.cu:58-87,130-145). */
int vx_fold_line_soa(const uint32_t *d_in0, const uint32_t *d_in1, const uint32_t *d_in2, const uint32_t *d_in3,
                     const uint32_t *d_twiddles, uint32_t *d_out0, uint32_t *d_out1, uint32_t *d_out2,
                     uint32_t *d_out3, const uint32_t *h_alpha, uint32_t half_n, vx_stream_t stream);
/* Replaces cuda_fold_circle_into_line_soa (ffi.rs:354-366, fri.cu:92-126,147-163). */
int vx_fold_circle_into_line_soa(uint32_t *d_dst0, uint32_t *d_dst1, uint32_t *d_dst2, uint32_t *d_dst3,
                                 const uint32_t *d_src0, const uint32_t *d_src1, const uint32_t *d_src2,
                                 const uint32_t *d_src3, const uint32_t *d_twiddles, const uint32_t *h_alpha,
                                 const uint32_t *h_alpha_sq, uint32_t half_n, vx_stream_t stream);

/* --------------------------------------------------------------- Merkle */

/* Replaces cuda_merkle_hash_leaves (ffi.rs:777-782, cuda/blake2s.cu:13-61,536-545;
 * callers src/merkle.rs:32-34,60,146-158): leaf = Blake2s(min(n_cols,16) LE
 * words). d_columns is a DEVICE array of device column pointers, as the
 * reference uploads it (a host array is accepted too). Reading the pointers
 * back waits for the stream's earlier work; the hashing is asynchronous. */
int vx_merkle_hash_leaves(const uint32_t *const *d_columns, uint32_t *d_hashes, uint32_t n_cols,
                          uint32_t n_leaves, vx_stream_t stream);
/* The same leaves from a HOST array of device column pointers, fully
 * asynchronous. No reference counterpart (the reference always uploads). */
int vx_merkle_hash_leaves_host(const uint32_t *const *h_columns, uint32_t *d_hashes, uint32_t n_cols,
                               uint32_t n_leaves, vx_stream_t stream);
/* Replace cuda_merkle_hash_leaves_forge_single / _quad (ffi.rs:797-815). */
int vx_merkle_hash_leaves_single(const uint32_t *d_col, uint32_t *d_hashes, uint32_t n_leaves, vx_stream_t stream);
int vx_merkle_hash_leaves_quad(const uint32_t *d_c0, const uint32_t *d_c1, const uint32_t *d_c2,
                               const uint32_t *d_c3, uint32_t *d_hashes, uint32_t n_leaves, vx_stream_t stream);
/* Replaces cuda_merkle_hash_nodes (ffi.rs:818-822, blake2s.cu:64-87,547-555). */
int vx_merkle_hash_nodes(const uint32_t *d_children, uint32_t *d_parents, uint32_t n_parents, vx_stream_t stream);
/* Replaces cuda_merkle_reduce_to_root (ffi.rs:824-828, blake2s.cu:244-297,558-565):
 * n_nodes a power of two <= 1024. */
int vx_merkle_reduce_to_root(const uint32_t *d_nodes, uint32_t *d_root, uint32_t n_nodes, vx_stream_t stream);
/* Replaces cuda_merkle_commit_small_soa4 (ffi.rs:830-835, blake2s.cu:92-189,569-580):
 * root of a QM31 SoA tree of n_leaves (power of two, 4..2048). */
int vx_merkle_commit_small_soa4(const uint32_t *d_c0, const uint32_t *d_c1, const uint32_t *d_c2,
                                const uint32_t *d_c3, uint32_t *d_root, uint32_t n_leaves, vx_stream_t stream);
/* Replaces cuda_merkle_hash_leaves_merge_soa4 (ffi.rs:837-842, blake2s.cu:195-240,584-596):
 * parents[j] = H(leaf(2j), leaf(2j+1)). */
int vx_merkle_hash_leaves_merge_soa4(const uint32_t *d_c0, const uint32_t *d_c1, const uint32_t *d_c2,
                                     const uint32_t *d_c3, uint32_t *d_parents, uint32_t n_leaves,
                                     vx_stream_t stream);
/* Replace cuda_merkle_tiled_soa4[_stream] / cuda_merkle_tiled_generic
 * (ffi.rs:844-864, blake2s.cu:305-532,600-644; callers src/merkle.rs:
 * 398-407): the root of every 1024-leaf tile, n_leaves any positive multiple
 * of 1024. vx_merkle_tiled_generic takes a DEVICE array of device column
 * pointers, as the reference does (a host array is accepted too; reading the
 * pointers back waits for the stream's earlier work). */
int vx_merkle_tiled_soa4(const uint32_t *d_c0, const uint32_t *d_c1, const uint32_t *d_c2, const uint32_t *d_c3,
                         uint32_t *d_subtree_roots, uint32_t n_leaves, vx_stream_t stream);
int vx_merkle_tiled_generic(const uint32_t *const *d_columns, uint32_t *d_subtree_roots, uint32_t n_cols,
                            uint32_t n_leaves, vx_stream_t stream);
/* The same tile roots from a HOST array of device column pointers, fully
 * asynchronous. No reference counterpart (the reference always uploads). */
int vx_merkle_tiled_generic_host(const uint32_t *const *h_columns, uint32_t *d_subtree_roots, uint32_t n_cols,
                                 uint32_t n_leaves, vx_stream_t stream);

/* ------------------------------------------------------------------ PoW */

/* Replaces cuda_grind_pow (ffi.rs:766-772, cuda/grind.cu:26-110): tests
 * nonces [batch_offset, batch_offset + n_threads) against the prefixed
 * digest and atomically lowers *d_result (caller initialises it to
 * UINT64_MAX). Unlike the reference, the result is the smallest valid nonce
 * of the batch (no racing early exit). */
int vx_grind_pow(const uint32_t *d_prefixed_digest, uint64_t *d_result, uint32_t pow_bits, uint64_t batch_offset,
                 uint32_t n_threads, vx_stream_t stream);
/* Whole deterministic grind: *d_result = smallest nonce >= 0 with
 * trailing_zeros(Blake2s(prefix || nonce)[0..16]) >= pow_bits
 * (Channel::verify_pow_nonce, src/channel.rs:140-168). pow_bits <= 32. */
int vx_grind_pow_min(const uint32_t *d_prefixed_digest, uint32_t pow_bits, uint64_t *d_result, vx_stream_t stream);
/* The stwo plugin's grind as one call (GrindOps for Blake2sChannelGeneric<false>,
 * crates/cuda-backend/src/grind_ops.rs:21-135, which loops on the host over
 * cuda_grind_pow batches of 2^20): *d_result = the smallest valid nonce of
 * the form (hi << 32) | low with low < 2^20, hi = 0, 1, ... (the order the
 * plugin and stwo's SIMD backend search); 0 when pow_bits = 0. pow_bits <= 32.
 * Async, except for pow_bits = 0. */
int vx_grind_pow_stwo(const uint32_t *d_prefixed_digest, uint32_t pow_bits, uint64_t *d_result, vx_stream_t stream);
/* Replaces cuda_grind_pow_m31_output (ffi.rs:1196-1204,
 * cuda/grind_m31_output.cu:15-87): vx_grind_pow with every word of the
 * Blake2s output reduced mod P before the trailing-zero count (stwo's
 * Blake2sChannelGeneric<true>). The prefixed digest is the caller's, already
 * reduced as the plugin does (grind_ops.rs:160-180). Smallest valid nonce of
 * the batch, as vx_grind_pow. */
int vx_grind_pow_m31_output(const uint32_t *d_prefixed_digest, uint64_t *d_result, uint32_t pow_bits,
                            uint64_t batch_offset, uint32_t n_threads, vx_stream_t stream);
/* The plugin's GrindOps for Blake2sChannelGeneric<true> as one call
 * (grind_ops.rs:154-245, which loops on the host over
 * cuda_grind_pow_m31_output batches of 2^20): vx_grind_pow_stwo's search
 * with the M31-reduced test. 0 when pow_bits = 0. pow_bits <= 32. */
int vx_grind_pow_stwo_m31(const uint32_t *d_prefixed_digest, uint32_t pow_bits, uint64_t *d_result,
                          vx_stream_t stream);

/* Poseidon252 (the Stark field p = 2^251 + 17 * 2^192 + 1). An element is
 * 4 little-endian u64 limbs of its Montgomery form a * 2^256 mod p
 * (starknet-ff's internal form, the plugin's CudaColumn<FieldElement252>). */

/* Replaces cuda_grind_pow_poseidon (ffi.rs:1205-1212,
 * cuda/grind_poseidon.cu:24-93): tests nonces [batch_offset, batch_offset +
 * n_threads) against poseidon_hash(prefixed, nonce), counting the trailing
 * zeros of the hash's low 128 bits, and atomically lowers *d_result (caller
 * initialises it to UINT64_MAX). d_prefixed_digest_mont = poseidon_hash_many(
 * [POW_PREFIX, digest, pow_bits]) in Montgomery form (grind_ops.rs:273-280).
 * Unlike the reference, the result is the smallest valid nonce of the batch
 * (no racing early exit). pow_bits <= 128. */
int vx_grind_pow_poseidon(const uint64_t *d_prefixed_digest_mont, uint64_t *d_result, uint32_t pow_bits,
                          uint64_t batch_offset, uint32_t n_threads, vx_stream_t stream);
/* The plugin's GrindOps for Poseidon252Channel as one call
 * (grind_ops.rs:263-348, which loops on the host over cuda_grind_pow_poseidon
 * batches of 2^20 from nonce 0): *d_result = the smallest valid nonce; 0 when
 * pow_bits = 0. pow_bits <= 40. Async, except for pow_bits = 0. */
int vx_grind_pow_stwo_poseidon(const uint64_t *d_prefixed_digest_mont, uint32_t pow_bits, uint64_t *d_result,
                               vx_stream_t stream);
/* Replaces build_leaves_poseidon252 (ffi.rs:948-963,
 * cuda/merkle_poseidon252.cu:37-123, caller merkle_ops.rs:247-311): leaf i =
 * poseidon_hash_many of the M31 values of every column at its lifted row
 * (row i for a column of 2^lifting_log_size rows, else
 * ((i >> (lifting_log_size - log + 1)) << 1) | (i & 1)), packed 8 to an
 * element, first value most significant, a short last group adding its
 * count times 2^248. d_col_ptrs and d_col_log_sizes are DEVICE arrays of
 * n_cols entries, as the plugin uploads them. n_leaves <= 2^lifting_log_size;
 * d_out holds 4 * n_leaves u64. */
int vx_build_leaves_poseidon252(const uint32_t *const *d_col_ptrs, const uint32_t *d_col_log_sizes, uint32_t n_cols,
                                uint32_t lifting_log_size, uint64_t *d_out, uint32_t n_leaves, vx_stream_t stream);
/* Replaces build_next_layer_poseidon252 (ffi.rs:965-974,
 * cuda/merkle_poseidon252.cu:127-148): parent i = poseidon_hash(prev[2i],
 * prev[2i + 1]) = permute([prev[2i], prev[2i + 1], 2])[0]. */
int vx_build_next_layer_poseidon252(const uint64_t *d_prev, uint64_t *d_out, uint32_t n_parents, vx_stream_t stream);

/* --------------------------------------------- prove / verify (src/prover.rs) */

typedef struct vx_prover vx_prover;

/* Size in bytes of a KRAK proof (v1 when pow_bits = 0, else v2). */
size_t vx_proof_size(uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits);

/* ProverCache::new counterpart (src/prover.rs:172-218): twiddles, all
 * intermediate buffers and the proof buffer for one size class.
 * stream NULL = a private non-blocking stream. */
int vx_prover_create(uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits,
                     vx_stream_t stream, vx_prover **out);
int vx_prover_destroy(vx_prover *p);
/* prove_cached counterpart (src/prover.rs:231-234): writes the KRAK bytes
 * (src/bin/stark_cli.rs:1092-1120) of the proof of Fib(a, b) to out. */
int vx_prover_prove(vx_prover *p, uint32_t a, uint32_t b, uint8_t *out, size_t out_cap, size_t *out_len);
/* Split form: enqueue all kernels without waiting, then wait + copy. */
int vx_prover_enqueue(vx_prover *p, uint32_t a, uint32_t b, int timed);
int vx_prover_wait(vx_prover *p, uint8_t *out, size_t out_cap, size_t *out_len);
/* Stage times (ms) of the last timed proof: trace, interpolate, lde,
 * trace_commit, quotient_commit, fri, queries, decommit, download, total,
 * and the trace-tree subtree hashing kernel alone. */
int vx_prover_stage_times(vx_prover *p, float out[11]);
/* Device bytes held by the prover (admission accounting). */
uint64_t vx_prover_device_bytes(const vx_prover *p);
/* Buffers of the last proof, for inspection: which = 0 trace/
 * coefficients (2^log_n), 1 LDE evaluation (2^log_eval), 2 + k FRI layer k
 * (4 SoA columns back to back), 100 the KRAK proof buffer, 101 the pinned
 * HOST copy of the last proof (valid after vx_prover_wait). */
int vx_prover_buffer(const vx_prover *p, int which, const uint32_t **d_ptr, uint64_t *words);

/* Synthetic multi-column AIR (the serving workload's 34-column request; no
 * reference counterpart, modelled on the reference's 34-column Cairo trace,
 * src/cairo_air/prover.rs:341-350): n_cols columns of seeded random M31
 * values, n_cons degree-2 transition constraints c t_a t_b' + d t_e with
 * coefficients from air_seed, trace committed in trees of <= 16 columns,
 * KRAK v3 bytes (checked by vx_verify). The proof is made by the same
 * vx_prover_* calls, with (a, b) = the trace seed's low and high words. */
int vx_d2_prover_create(uint32_t log_n, uint32_t n_cols, uint32_t n_cons, uint64_t air_seed, uint32_t blowup_bits,
                        uint32_t n_queries, uint32_t pow_bits, vx_stream_t stream, vx_prover **out);
size_t vx_d2_proof_size(uint32_t log_n, uint32_t n_cols, uint32_t blowup_bits, uint32_t n_queries);

/* Batched-columns proof (KRAK v4; the batched uneven-commit workload as a
 * full proof): n_cols columns of log sizes h_log_sizes (sorted ascending,
 * 14..24) with values from the synthetic generator, bit-reversed LDE, one
 * lifted tree (cuda/merkle_leaves_lifted.cu), a random-linear-combination
 * quotient (the transcript mixes Blake2s of the log sizes, then the lifted
 * root, then draws the combination challenge), then the Fibonacci FRI / PoW
 * / queries. Proved by vx_prover_*
 * with (a, b) = the trace seed's low and high words; checked by vx_verify. */
int vx_batch_prover_create(const uint32_t *h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t n_queries,
                           uint32_t pow_bits, vx_stream_t stream, vx_prover **out);
size_t vx_batch_proof_size(const uint32_t *h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t n_queries);

/* Untimed proofs replay a HIP graph of the whole proof captured on first
 * use (default on); 0 launches the kernels one by one instead. Same bytes.
 * No counterpart in the reference (its prove_cached launches per call). */
int vx_prover_set_graphs(vx_prover *p, int on);

/* prover::prove counterpart (src/prover.rs:221-223): one-shot proof. */
int vx_prove(uint32_t a, uint32_t b, uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits,
             uint8_t *out, size_t out_cap, size_t *out_len);
/* verifier::verify counterpart (src/verifier.rs:39-236) over KRAK bytes:
 * VX_OK if accepted, VX_ERR_VERIFY (reason in err) if rejected. Host only. */
int vx_verify(const uint8_t *proof, size_t len, uint32_t blowup_bits, uint32_t n_queries, char *err,
              size_t err_cap);

/* ------------------------------------------------------------ stwo export
 * Host only. cairo_proof_to_json counterpart (src/stwo_export.rs:401-676)
 * for the KRAK proofs of this library (v1..v4): the stwo-shaped JSON
 * (config, tree_meta, commitments, sampled_values, decommitments with
 * deduplicated hash_witness, queried_values[tree][col][query], query_indices,
 * proof_of_work, fri_proof with fri_witness and the bit-reversed LinePoly
 * coefficients of the last layer), plus a "krak_openings" object with the
 * leaves each hash witness proves. *out_len = the JSON length; out may be
 * NULL to query it, else cap must exceed it (a NUL is appended). blowup_bits
 * is read from the proof except for v1, which does not carry it. */
int vx_export_stwo_json(const uint8_t *proof, size_t len, uint32_t blowup_bits, char *out, size_t cap,
                        size_t *out_len);
/* auth_paths_to_hash_witness (src/stwo_export.rs:262-320): positions[n],
 * paths = n auth paths of `height` 8-word hashes (leaf level first) back to
 * back. *out_n = witness hashes; out (8 words each) may be NULL. */
int vx_stwo_hash_witness(const uint64_t *positions, size_t n, const uint32_t *paths, uint32_t height,
                         uint32_t *out, size_t cap, size_t *out_n);
/* dedup_fri_witness (src/stwo_export.rs:174-206): sibling_values[n] QM31
 * (4 words each) at positions[i] ^ 1. *out_n = witness values; out may be NULL. */
int vx_stwo_fri_witness(const uint64_t *positions, size_t n, const uint32_t *sibling_values, uint32_t *out,
                        size_t cap, size_t *out_n);
/* reconstruct_root_from_witness (src/stwo_export.rs:697-780): leaves are
 * Blake2s of the n_cols values (little-endian words) at each position,
 * col_values column-major (n_cols x n). VX_ERR_VERIFY with the reason in err
 * when the witness is short, long or leaves more than one root. */
int vx_stwo_reconstruct_root(uint32_t log_height, const uint64_t *positions, size_t n, const uint32_t *col_values,
                             uint32_t n_cols, const uint32_t *witness, size_t n_witness, uint32_t root[8], char *err,
                             size_t err_cap);
/* last_layer_poly_coeffs (src/fri.rs:294-335) followed by the bit-reversal
 * of src/cairo_air/prover.rs:2896-2900: n QM31 last-layer evaluations
 * (bit-reversed order on the half_odds line domain) -> n bit-reversed
 * LinePoly coefficients. */
int vx_stwo_last_layer_coeffs(const uint32_t *evals, uint32_t n, uint32_t *out);

/* ------------------------------------------------- batched uneven columns
 * Many M31 columns of different log sizes through one launch per phase,
 * driven by device work lists. The reference batches only equal-size
 * columns, one launch per layer (src/ntt.rs:191-281,
 * cuda/circle_ntt.cu:314-404). Twiddle tables are built by the library per
 * distinct size (shared with provers of that size). */

/* Synthetic data: d[i] = mix64(seed + G ((col << 32 | i) + 1)) mod P for
 * i < 2^log_n (G = 0x9E3779B97F4A7C15, mix64 = the SplitMix64 finaliser):
 * column `col` of the synthetic AIR's trace generator. Synchronises. */
int vx_m31_fill_random(uint32_t *d_data, uint32_t log_n, uint64_t seed, uint32_t col, vx_stream_t stream);

/* Interpolate column c (2^log_sizes[c] values at h_in[c], 14..24) on
 * half_coset(log), write the coefficients to h_coef[c] and the evaluation on
 * the 2^blowup-times larger eval domain to h_eval[c]: the same values as
 * vx_circle_ntt_interpolate + vx_circle_ntt_evaluate of the zero-padded
 * coefficients (src/prover.rs:293-318), per column. Host arrays of device
 * pointers; synchronises the stream. */
int vx_circle_ntt_lde_batch(const uint32_t *const *h_in, uint32_t *const *h_coef, uint32_t *const *h_eval,
                            const uint32_t *h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, vx_stream_t stream);
/* Out-of-place dst[c][bitrev(i)] = src[c][i], log sizes >= 10
 * (cuda_bit_reverse_m31, src/cuda/ffi.rs:311, for many columns at once). */
int vx_bit_reverse_m31_batch(const uint32_t *const *h_src, uint32_t *const *h_dst, const uint32_t *h_log_sizes,
                             uint32_t n_cols, vx_stream_t stream);
/* cuda_build_leaves_lifted (src/cuda/ffi.rs:937-947,
 * cuda/merkle_leaves_lifted.cu): leaf i of a 2^lifting_log_size tree =
 * Blake2s of every column's value at its lifted row
 * ((i >> (S - L + 1)) << 1 | (i & 1) for a column of log size L < S), columns
 * sorted by ascending size. The schedule of same-size chunks is derived from
 * h_log_sizes. d_hashes: 8 words per leaf. */
int vx_build_leaves_lifted(const uint32_t *const *h_cols, const uint32_t *h_log_sizes, uint32_t n_cols,
                           uint32_t lifting_log_size, uint32_t *d_hashes, vx_stream_t stream);
/* Root of that lifted tree (lifting size = the largest column), nodes
 * Blake2s(left || right). */
int vx_merkle_root_lifted(const uint32_t *const *h_cols, const uint32_t *h_log_sizes, uint32_t n_cols,
                          uint32_t *d_root, vx_stream_t stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same root as vx_merkle_root_lifted, computed with
 * no device allocation. The column table and the stored tree levels live in
 * a static 8 MiB device scratch array per GPU (held until the process ends);
 * shapes that need more (a table of 16 B per column plus about
 * 64 B x 2^(log_max - 10) of tree levels, i.e. log_max above 26) are rejected
 * with VX_ERR_ARG. At most 4096 columns. Asynchronous on `stream`: the root
 * is in d_root once the stream reaches it. Every call on one GPU uses the
 * same scratch, so calls that overlap on one GPU race: serialise them. */
int vx_merkle_root_lifted_noalloc(const uint32_t *const *h_cols, const uint32_t *h_log_sizes, uint32_t n_cols,
                                  uint32_t *d_root, vx_stream_t stream);

/* ------------------------------------------------------ Cairo VM AIR quotient */

/* Replaces cuda_cairo_quotient (src/cuda/ffi.rs:592-609, cuda/cairo_constraint.cu:
 * 157-580, caller src/cairo_air/prover.rs:2046-2066). Q[i] = (sum of the 40
 * alpha-weighted constraints of row i) * vh_inv[i] into out0..out3 (QM31 SoA),
 * i < n; the next row is (i + blowup_step) mod n. d_trace_cols is a DEVICE
 * array of the 34 trace-column pointers (pc, ap, fp, inst_lo, inst_hi, 15
 * flags, dst_addr, dst, op0_addr, op0, op1_addr, op1, res, off0..2, dst_inv,
 * dict key / new / active), read back after the stream's earlier work.
 * Interaction columns: S_logup, T1, T2, T3, S_rc, U1, U2, S_dict (4 SoA words
 * each). d_alpha: 40 QM31; d_challenges: z_mem, alpha_mem, alpha_mem^2, z_rc,
 * z_dict, alpha_dict (24 words). Asynchronous. */
int vx_cairo_quotient(const uint32_t *const *d_trace_cols, const uint32_t *s_logup0, const uint32_t *s_logup1,
                      const uint32_t *s_logup2, const uint32_t *s_logup3, const uint32_t *t1l0, const uint32_t *t1l1,
                      const uint32_t *t1l2, const uint32_t *t1l3, const uint32_t *t2l0, const uint32_t *t2l1,
                      const uint32_t *t2l2, const uint32_t *t2l3, const uint32_t *t3l0, const uint32_t *t3l1,
                      const uint32_t *t3l2, const uint32_t *t3l3, const uint32_t *s_rc0, const uint32_t *s_rc1,
                      const uint32_t *s_rc2, const uint32_t *s_rc3, const uint32_t *u1r0, const uint32_t *u1r1,
                      const uint32_t *u1r2, const uint32_t *u1r3, const uint32_t *u2r0, const uint32_t *u2r1,
                      const uint32_t *u2r2, const uint32_t *u2r3, const uint32_t *s_dict0, const uint32_t *s_dict1,
                      const uint32_t *s_dict2, const uint32_t *s_dict3, uint32_t *out0, uint32_t *out1,
                      uint32_t *out2, uint32_t *out3, const uint32_t *d_alpha, const uint32_t *d_vh_inv,
                      const uint32_t *d_trans_factor, const uint32_t *d_challenges, uint32_t n,
                      uint32_t blowup_step, vx_stream_t stream);
/* Replaces cuda_cairo_quotient_slab (ffi.rs:611-638): the columns are slabs
 * of chunk_n + blowup_step rows starting at global row chunk_offset (next
 * row = local row + blowup_step); vh_inv, trans_factor and out are global
 * and indexed by chunk_offset + local row. */
int vx_cairo_quotient_slab(const uint32_t *const *d_trace_cols, const uint32_t *s_logup0, const uint32_t *s_logup1,
                           const uint32_t *s_logup2, const uint32_t *s_logup3, const uint32_t *t1l0,
                           const uint32_t *t1l1, const uint32_t *t1l2, const uint32_t *t1l3, const uint32_t *t2l0,
                           const uint32_t *t2l1, const uint32_t *t2l2, const uint32_t *t2l3, const uint32_t *t3l0,
                           const uint32_t *t3l1, const uint32_t *t3l2, const uint32_t *t3l3, const uint32_t *s_rc0,
                           const uint32_t *s_rc1, const uint32_t *s_rc2, const uint32_t *s_rc3, const uint32_t *u1r0,
                           const uint32_t *u1r1, const uint32_t *u1r2, const uint32_t *u1r3, const uint32_t *u2r0,
                           const uint32_t *u2r1, const uint32_t *u2r2, const uint32_t *u2r3,
                           const uint32_t *s_dict0, const uint32_t *s_dict1, const uint32_t *s_dict2,
                           const uint32_t *s_dict3, uint32_t *out0, uint32_t *out1, uint32_t *out2, uint32_t *out3,
                           const uint32_t *d_alpha, const uint32_t *d_vh_inv, const uint32_t *d_trans_factor,
                           const uint32_t *d_challenges, uint32_t chunk_n, uint32_t blowup_step,
                           uint32_t chunk_offset, vx_stream_t stream);
/* Replaces cuda_cairo_quotient_chunk (ffi.rs:640-655, cairo_constraint.cu:
 * 582-850): rows offset .. offset + chunk_n of global_n in the chunked form
 * (35 weighted constraints; the LogUp and range-check steps as S_next - S
 * minus the sum of inverse denominators), output at the chunk-local row. */
int vx_cairo_quotient_chunk(const uint32_t *const *d_trace_cols, const uint32_t *s_logup0,
                            const uint32_t *s_logup1, const uint32_t *s_logup2, const uint32_t *s_logup3,
                            const uint32_t *s_rc0, const uint32_t *s_rc1, const uint32_t *s_rc2,
                            const uint32_t *s_rc3, const uint32_t *s_dict0, const uint32_t *s_dict1,
                            const uint32_t *s_dict2, const uint32_t *s_dict3, uint32_t *out0, uint32_t *out1,
                            uint32_t *out2, uint32_t *out3, const uint32_t *d_alpha, const uint32_t *d_vh_inv,
                            const uint32_t *d_trans_factor, const uint32_t *d_challenges, uint32_t offset,
                            uint32_t chunk_n, uint32_t global_n, uint32_t blowup_step, vx_stream_t stream);
/* Fused form of vx_cairo_quotient followed by the commitment of its four
 * output columns (the reference runs cuda_cairo_quotient, then commits the
 * quotient columns, src/cairo_air/prover.rs:2046-2066 and the quotient
 * commitment after it): one subtree launch computes Q of rows 0..2^log_n
 * (next row (i + blowup_step) mod 2^log_n), stores it in out0..out3
 * (16-byte aligned) and hashes QM31 leaves Blake2s(4 words); tree levels
 * store_lo..log_n land in d_levels (vx_merkle_levels_words(log_n, store_lo)
 * words), the root in d_root. h_inter is a HOST array of the 32 interaction
 * column pointers in vx_cairo_quotient's order. Synchronous. */
int vx_cairo_quotient_commit(const uint32_t *const *d_trace_cols, const uint32_t *const *h_inter, uint32_t *out0,
                             uint32_t *out1, uint32_t *out2, uint32_t *out3, const uint32_t *d_alpha,
                             const uint32_t *d_vh_inv, const uint32_t *d_trans_factor, const uint32_t *d_challenges,
                             uint32_t log_n, uint32_t blowup_step, uint32_t store_lo, uint32_t *d_levels,
                             uint32_t *d_root, vx_stream_t stream);
/* Replace cuda_compute_vanishing_inv / cuda_compute_vanishing (ffi.rs:570-586,
 * cairo_constraint.cu:77-118,990-1019): Z_H of half_coset(log_n) (or its
 * inverse, 0 for 0) at the point initial * step^bitrev(i) for i < 2^log_eval. */
int vx_compute_vanishing_inv(uint32_t initial_x, uint32_t initial_y, uint32_t step_x, uint32_t step_y,
                             uint32_t *d_out, uint32_t log_eval, uint32_t log_n, vx_stream_t stream);
int vx_compute_vanishing(uint32_t initial_x, uint32_t initial_y, uint32_t step_x, uint32_t step_y, uint32_t *d_out,
                         uint32_t log_eval, uint32_t log_n, vx_stream_t stream);
/* Replaces cuda_axpy_m31 (ffi.rs:588-590, cairo_constraint.cu:1037-1057): y += scalar * x. */
int vx_axpy_m31(uint32_t scalar, const uint32_t *d_x, uint32_t *d_y, uint32_t n, vx_stream_t stream);

/* ------------------------------------------------ bytecode constraint eval
 * Replaces cuda_bytecode_constraint_eval (ffi.rs:886-902,
 * cuda/constraint_eval.cu:97-451; called from the stwo backend's
 * component_prover.rs:478-510): the register bytecode of
 * crates/cuda-backend/src/constraint_eval/bytecode.rs (device pointer,
 * n_words words) run for every row of the evaluation domain (n_rows, a
 * power of two, bit-reversed order) over the device array d_trace_cols of
 * n_trace_cols column pointers (a null column reads as 0); accum_c[row] +=
 * (denom_inv[row >> log2(trace_n_rows)] * sum_k coeff_k * constraint_k)_c
 * with coeff_k the k-th QM31 of d_random_coeff_powers. trace_col_sizes,
 * log_expand and n_registers are accepted and unused (the first two are
 * implied, the register file is sized from the program). Unlike the
 * reference, a malformed program (unknown opcode, truncated instruction,
 * column out of range, a register read before it is written, a constant
 * that is not a canonical M31 value, row offsets with n_rows == trace_n_rows)
 * is rejected with VX_ERR_ARG and nothing is written; any 16-bit register
 * index is accepted. The program is read back to the host, so the call
 * synchronises `stream` before and after the evaluation. */
int vx_bytecode_constraint_eval(const uint32_t *d_bytecode, uint32_t n_words, const uint32_t *const *d_trace_cols,
                                const uint32_t *d_trace_col_sizes, uint32_t n_trace_cols, uint32_t n_rows,
                                uint32_t trace_n_rows, const uint32_t *d_random_coeff_powers,
                                const uint32_t *d_denom_inv, uint32_t log_expand, uint32_t *accum0,
                                uint32_t *accum1, uint32_t *accum2, uint32_t *accum3, uint32_t n_registers,
                                vx_stream_t stream);
/* Replaces cuda_warp_bytecode_constraint_eval (ffi.rs:908-925,
 * cuda/constraint_eval_warp.cu), the reference's kernel for programs with
 * more than 1,024 registers: the same evaluation as above (registers are
 * renamed to the program's live values, so no second kernel is needed). */
int vx_warp_bytecode_constraint_eval(const uint32_t *d_bytecode, uint32_t n_words,
                                     const uint32_t *const *d_trace_cols, const uint32_t *d_trace_col_sizes,
                                     uint32_t n_trace_cols, uint32_t n_rows, uint32_t trace_n_rows,
                                     const uint32_t *d_random_coeff_powers, const uint32_t *d_denom_inv,
                                     uint32_t log_expand, uint32_t *accum0, uint32_t *accum1, uint32_t *accum2,
                                     uint32_t *accum3, uint32_t n_registers, vx_stream_t stream);
/* No reference counterpart: by default a program is compiled for gfx950 on
 * first use (hipRTC; straight-line code, values in VGPRs; cached per device,
 * program and domain shape) and programs of more than 50,000 instructions
 * are interpreted; on = 0 interprets every program (same results). */
int vx_bytecode_set_jit(int on);
/* Host only, no reference counterpart: the HIP source the compiled tier
 * generates for a program (host pointer) over 2^eval_log rows and a trace
 * domain of 2^trace_log (out may be NULL to query *out_len), and a compile
 * of it without loading (*code_bytes = the gfx950 code object's size). */
int vx_bytecode_jit_source(const uint32_t *h_bytecode, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                           uint32_t trace_log, char *out, size_t cap, size_t *out_len);
int vx_bytecode_jit_compile(const uint32_t *h_bytecode, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                            uint32_t trace_log, size_t *code_bytes);
/* Host only, no reference counterpart: checks a program (host pointer) as
 * the evaluation does and reports its register file after renaming (M31
 * slots of one word per row, QM31 slots of four) and its constraint count. */
int vx_bytecode_slots(const uint32_t *h_bytecode, uint32_t n_words, uint32_t n_trace_cols, uint32_t *m31_slots,
                      uint32_t *qm31_slots, uint32_t *n_constraints);

/* Resident batch (the batched-commit workload): inputs, coefficients and
 * evaluations of n_cols columns (log sizes sorted ascending, 14..24) kept in
 * HBM; each run is four work-list LDE launches, one bit-reverse launch and one
 * lifted commitment. */
typedef struct vx_batch vx_batch;
int vx_batch_create(const uint32_t *h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, int bit_reverse,
                    vx_stream_t stream, vx_batch **out);
int vx_batch_destroy(vx_batch *b);
/* which: 0 input values (2^L), 1 coefficients (2^L), 2 evaluation
 * (2^(L+blowup), bit-reversed when enabled). */
int vx_batch_column(const vx_batch *b, uint32_t col, int which, uint32_t **d_ptr, uint64_t *words);
int vx_batch_enqueue(vx_batch *b, int timed);
/* Wait for the run; h_root (host, may be NULL) receives the lifted root. */
int vx_batch_wait(vx_batch *b, uint32_t h_root[8]);
/* ms of the last timed run: lde, bit reverse, commit, total. */
int vx_batch_stage_times(vx_batch *b, float out[4]);
uint64_t vx_batch_device_bytes(const vx_batch *b);

/* ------------------------------------------------------------------ serving
 * Replaces the reference's one-worker proof queue (src/bin/dashboard.rs:
 * 66-96, drained one proof at a time; ProverPipeline::prove_batch,
 * src/prover.rs:2278-2319, is the same serial loop) with continuous
 * batching: n_slots HIP streams, each request starts on a free slot as soon
 * as one frees up and overlaps on the device with the others. Each slot keeps
 * a resident prover per size class, admitted against mem_budget_bytes of
 * device memory (idle provers are evicted least-recently-used). n_slots = 1
 * is the one-at-a-time loop. Proof bytes are identical to vx_prover_prove. */
typedef struct vx_server vx_server;
typedef struct vx_request {
  uint32_t air; /* 0 = Fibonacci, 1 = synthetic multi-column AIR */
  uint32_t log_n, blowup_bits, n_queries, pow_bits;
  uint32_t a, b;           /* Fibonacci public inputs */
  uint32_t n_cols, n_cons; /* air 1: trace columns, constraints */
  uint32_t reserved;
  uint64_t seed;           /* air 1: trace seed */
  uint64_t air_seed;       /* air 1: constraint coefficient seed */
} vx_request;

int vx_server_create(uint32_t n_slots, uint64_t mem_budget_bytes, int use_graphs, vx_server **out);
int vx_server_destroy(vx_server *s);
/* Create the request's size class on every slot ahead of traffic. */
int vx_server_prepare(vx_server *s, const vx_request *r);
/* Queue a request (non-blocking); out (host, out_cap >= vx_proof_size) receives
 * the proof before the ticket completes. */
int vx_server_submit(vx_server *s, const vx_request *r, uint8_t *out, size_t out_cap, uint64_t *ticket);
/* Block until the ticket completes and release it; times = submit, start
 * and completion (seconds on the server's clock). */
int vx_server_wait(vx_server *s, uint64_t ticket, size_t *out_len, double times[3]);
/* 1 if the ticket has completed, 0 if not, < 0 on error. */
int vx_server_done(vx_server *s, uint64_t ticket);
uint64_t vx_server_resident_bytes(const vx_server *s);
/* Scheduler counters: out[0] provers created, out[1] provers evicted,
 * out[2] admissions postponed until a running proof or a proof copy-out
 * released memory. No reference counterpart (the reference keeps no
 * resident provers). */
int vx_server_stats(const vx_server *s, uint64_t out[3]);
/* Replay an arrival trace: request i is submitted arrival_s[i] seconds
 * (sorted) after the call starts; returns when all completed. times[3i..3i+2]
 * = submit, start, completion, in seconds from the call's start. */
int vx_server_run_trace(vx_server *s, const vx_request *reqs, const double *arrival_s, uint32_t n,
                        uint8_t *const *outs, const size_t *caps, size_t *lens, double *times);

/* --------------------------------------------------------- fused commits
 * The fused entry points named by the boundary plan: one call hashes every
 * leaf (producing the leaf words on the fly where noted), stores tree levels
 * store_lo..log_n into d_levels (level l at word offset
 * 8 * sum_{store_lo <= k < l} 2^(log_n - k), 8 words per node; size from
 * vx_merkle_levels_words) and writes the root (8 words) to d_root. They
 * replace the reference's leaf-hash / tiled-subtree / reduce-to-root /
 * host-upper-tree sequence (src/merkle.rs:236-451, cuda/blake2s.cu:305-532).
 * Alphas are host [4] words, as in cuda_fold_line_soa. The calls that take
 * alphas return after their work completes. */
uint64_t vx_merkle_levels_words(uint32_t log_n, uint32_t store_lo);
/* MerkleTree::commit_root_only_with_subtrees (src/merkle.rs:386-451) over
 * 1..16 columns of 2^log_n rows (leaf = the row's words). */
int vx_merkle_commit_tiled(const uint32_t *const *h_cols, uint32_t n_cols, uint32_t log_n, uint32_t store_lo,
                           uint32_t *d_levels, uint32_t *d_root, vx_stream_t stream);
/* cuda_fibonacci_quotient + MerkleTree::commit_root_soa4_with_subtrees
 * (src/prover.rs:341-365): the QM31 quotient q_i = alpha (e[i+2] - e[i+1] -
 * e[i]) of a 2^log_n evaluation, committed without being stored. */
int vx_quotient_commit(const uint32_t *d_eval, const uint32_t *alpha, uint32_t log_n, uint32_t store_lo,
                       uint32_t *d_levels, uint32_t *d_root, vx_stream_t stream);
/* fold_circle_into_line_with_twiddles of that quotient (dst = 0) + the
 * layer's commitment (src/prover.rs:388-413): writes line layer 0
 * (2^(log_n-1) QM31, SoA o0..o3) and commits it (tree of log_n - 1 levels). */
int vx_circle_fold_commit(const uint32_t *d_eval, const uint32_t *alpha_q, const uint32_t *alpha_f,
                          const uint32_t *d_twiddles, uint32_t *o0, uint32_t *o1, uint32_t *o2, uint32_t *o3,
                          uint32_t log_n, uint32_t store_lo, uint32_t *d_levels, uint32_t *d_root,
                          vx_stream_t stream);
/* fold_line_with_twiddles + the new layer's commitment (src/prover.rs:
 * 452-493): folds 2^(log_half_n + 1) SoA values into 2^log_half_n and
 * commits them. */
int vx_fold_line_commit(const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, const uint32_t *i3,
                        const uint32_t *d_twiddles, uint32_t *o0, uint32_t *o1, uint32_t *o2, uint32_t *o3,
                        const uint32_t *alpha, uint32_t log_half_n, uint32_t store_lo, uint32_t *d_levels,
                        uint32_t *d_root, vx_stream_t stream);

/* ------------------------------------------------------------ autotuner
 * No counterpart in the reference. The first commitment of each key (leaf
 * kind, columns, constraints, log size, stored levels) on a device is timed
 * with every valid subtree shape (leaves per workgroup and per lane) and the
 * fastest is cached; all shapes give the same tree, so proofs do not change.
 * Provers run one eager proof before capturing their graph so that every
 * commitment of the class is tuned. enabled = 0 uses the built-in shapes.
 * cache_file (may be NULL): entries are loaded from it and appended to it. */
int vx_tuner_configure(int enabled, const char *cache_file);
/* Up to `capacity` cached entries, 9 words each: device, leaf kind, n_cols,
 * n_cons, lifted columns, log_n, stored-from level, tile log, leaves-per-lane
 * log; *n = the number of entries (may exceed capacity). */
int vx_tuner_entries(uint32_t *out, uint32_t capacity, uint32_t *n);
/* The circle NTT's tuned plans (no counterpart in the reference, whose kernel
 * choices are compile-time features, Cargo.toml:30-52): the first transform
 * of each shape (direction, log_n, transformed bits, columns, basis) on a
 * device is timed with the built-in plan, with tiles in launch order instead
 * of XCD-contiguous ranges, and with the strided bits over one more pass;
 * the fastest is cached (the built-in one unless another is more than 3 %
 * faster) and every plan computes the same values. Up to `capacity` entries,
 * 8 words each: device, forward, log_n, transformed bits, n_cols, stwo basis,
 * split, order; *n = the number of entries. Cache-file lines
 * "ntt fwd log_n bits n_cols stwo split order" are loaded by
 * vx_tuner_configure. */
int vx_tuner_ntt_entries(uint32_t *out, uint32_t capacity, uint32_t *n);

/* ------------------------------------------------ GKR (stwo plugin GkrOps)
 * The plugin's GKR sum-check operations (crates/cuda-backend/src/gkr_ops.rs
 * over cuda/gkr.cu). Secure (QM31) columns hold element i in words
 * 4i..4i+3, base columns one M31 word per element; every pointer is a device
 * address and every call is asynchronous on `stream`. n is the input length
 * in elements (even, >= 2). QM31 arguments (r, v, y_i, lambda) are 4 words
 * in device memory, as the plugin uploads them. */
/* Replaces cuda_gkr_fix_first_variable_base (ffi.rs:1129-1133): out[i] =
 * r (in[i + n/2] - in[i]) + in[i] for i < n/2, M31 in, QM31 out. */
int vx_gkr_fix_first_variable_base(const uint32_t *d_in, uint32_t *d_out, const uint32_t *d_r, uint32_t n,
                                   vx_stream_t stream);
/* Replaces cuda_gkr_fix_first_variable_secure (ffi.rs:1134-1138): the same
 * over a QM31 column; d_out may equal d_in. */
int vx_gkr_fix_first_variable_secure(const uint32_t *d_in, uint32_t *d_out, const uint32_t *d_r, uint32_t n,
                                     vx_stream_t stream);
/* Replaces cuda_gkr_gen_eq_evals_init (ffi.rs:1139-1140): buf[0] = v. */
int vx_gkr_gen_eq_evals_init(uint32_t *d_buf, const uint32_t *d_v, vx_stream_t stream);
/* Replaces cuda_gkr_gen_eq_evals_pass (ffi.rs:1141-1142): for j < cur_size,
 * t = buf[j] y_i, buf[cur_size + j] = t, buf[j] -= t. */
int vx_gkr_gen_eq_evals_pass(uint32_t *d_buf, const uint32_t *d_y_i, uint32_t cur_size, vx_stream_t stream);
/* No counterpart in the reference, whose gen_eq_evals (gkr_ops.rs:129-161)
 * calls init and then k passes with y[k-1], ..., y[0] from the host: the
 * same 2^k entries (v times the eq polynomial of y, y[0] on the index's top
 * bit) in one launch. d_y: k QM31 values; k <= 31. */
int vx_gkr_gen_eq_evals(uint32_t *d_buf, const uint32_t *d_y, uint32_t k, const uint32_t *d_v, vx_stream_t stream);
/* Replace cuda_gkr_next_layer_grand_product / _logup_generic /
 * _logup_mult / _logup_singles (ffi.rs:1143-1159): output element i from
 * the input pair (2i, 2i + 1); the product, or the fraction sum
 * (n0 d1 + n1 d0, d0 d1) with QM31 numerators, M31 numerators (one word
 * each) or unit numerators (no d_in_num). n input elements, n/2 out. */
int vx_gkr_next_layer_grand_product(const uint32_t *d_in, uint32_t *d_out, uint32_t n, vx_stream_t stream);
int vx_gkr_next_layer_logup_generic(const uint32_t *d_in_num, const uint32_t *d_in_den, uint32_t *d_out_num,
                                    uint32_t *d_out_den, uint32_t n, vx_stream_t stream);
int vx_gkr_next_layer_logup_mult(const uint32_t *d_in_num, const uint32_t *d_in_den, uint32_t *d_out_num,
                                 uint32_t *d_out_den, uint32_t n, vx_stream_t stream);
int vx_gkr_next_layer_logup_singles(const uint32_t *d_in_den, uint32_t *d_out_num, uint32_t *d_out_den, uint32_t n,
                                    vx_stream_t stream);
/* Replace cuda_gkr_sum_poly_grand_product / _logup_generic / _logup_mult /
 * _logup_singles (ffi.rs:1160-1179), the device part of
 * sum_as_poly_in_first_variable (gkr_ops.rs:256-330): with the pair of term
 * i at t = 0 being elements (2i, 2i + 1), at t = 1 (2(n_terms + i),
 * 2(n_terms + i) + 1) and at t = 2 twice t = 1 minus t = 0, writes
 * (sum_i eq[i] F(t = 0), sum_i eq[i] F(t = 2)) as 8 words per block to
 * d_partial_sums and the number of blocks to *n_blocks (the reference
 * returns it); F is the pair's product or its fraction sum plus lambda
 * times its denominator. Here every partial is folded into the first block,
 * so *n_blocks = 1; d_partial_sums must hold ceil(n_terms / 256) blocks,
 * as the plugin allocates. */
int vx_gkr_sum_poly_grand_product(const uint32_t *d_eq, const uint32_t *d_layer, uint32_t *d_partial_sums,
                                  uint32_t n_terms, uint32_t *n_blocks, vx_stream_t stream);
int vx_gkr_sum_poly_logup_generic(const uint32_t *d_eq, const uint32_t *d_in_num, const uint32_t *d_in_den,
                                  const uint32_t *d_lambda, uint32_t *d_partial_sums, uint32_t n_terms,
                                  uint32_t *n_blocks, vx_stream_t stream);
int vx_gkr_sum_poly_logup_mult(const uint32_t *d_eq, const uint32_t *d_in_num, const uint32_t *d_in_den,
                               const uint32_t *d_lambda, uint32_t *d_partial_sums, uint32_t n_terms,
                               uint32_t *n_blocks, vx_stream_t stream);
int vx_gkr_sum_poly_logup_singles(const uint32_t *d_eq, const uint32_t *d_in_den, const uint32_t *d_lambda,
                                  uint32_t *d_partial_sums, uint32_t n_terms, uint32_t *n_blocks,
                                  vx_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* VORTEX_HIP_H */
