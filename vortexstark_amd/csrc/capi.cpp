// extern "C" entry points of include/vortex_hip.h. Each shim validates its
// arguments on the host (shapes that the kernels assume), launches on the
// given stream and converts exceptions into a status code + vx_last_error().
#include "../../include/vortex_hip.h"

#include <cstring>
#include <string>

#include "prover.h"
#include "batch.h"
#include "export.h"
#include "vx_bytecode.h"
#include "tune.h"
#include "server.h"
#include "vx_cairo.h"
#include "vx_d2.h"

using namespace vx;

namespace vx {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace vx

template <class F>
static int guard(F&& f) {
  try {
    f();
    return VX_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return VX_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return VX_ERR_INTERNAL;
  }
}

static hipStream_t S(vx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }
static bool pow2(uint64_t n) { return n && (n & (n - 1)) == 0; }
static uint32_t log2u(uint64_t n) {
  uint32_t l = 0;
  while ((1ull << l) < n) l++;
  return l;
}
static qm31 host_q(const uint32_t* a) {
  require(a != nullptr, "alpha must be a host [4] array");
  for (int i = 0; i < 4; i++) require(a[i] < P, "alpha words must be canonical M31");
  return {{a[0], a[1], a[2], a[3]}};
}

// Reference twiddle tables for a transform of n = 2^log_n points.
static TwiddleRef tw_ref(const uint32_t* line, const uint32_t* circle, const uint32_t* h_offsets,
                         uint32_t n_line_layers, uint32_t n) {
  require(line && circle && pow2(n) && n >= 2, "ntt: n must be a power of two >= 2 with twiddles");
  uint32_t log_n = log2u(n);
  require(n_line_layers == log_n - 1, "ntt: n_line_layers must be log2(n) - 1");
  require(n_line_layers == 0 || h_offsets != nullptr, "ntt: null layer offsets");
  TwiddleRef r{};
  r.line = line;
  r.circle = circle;
  r.log_n = log_n;
  for (uint32_t l = 0; l < n_line_layers; l++) r.off[l] = h_offsets[l];
  return r;
}

extern "C" {

const char* vx_last_error(void) { return g_last_error.c_str(); }

int vx_device_sync(void) {
  return guard([&] { VX_HIP(hipDeviceSynchronize()); });
}

int vx_malloc(void** d, size_t bytes) {
  return guard([&] {
    require(d != nullptr, "malloc: null out");
    *d = nullptr;
    VX_HIP(hipMalloc(d, bytes));
  });
}

int vx_free(void* d) {
  return guard([&] { VX_HIP(hipFree(d)); });
}

static hipMemcpyKind copy_kind(int kind) {
  require(kind >= 1 && kind <= 3, "memcpy: kind must be 1 (H2D), 2 (D2H) or 3 (D2D)");
  return static_cast<hipMemcpyKind>(kind);
}

int vx_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  return guard([&] {
    require((dst && src) || bytes == 0, "memcpy: null pointer");
    VX_HIP(hipMemcpy(dst, src, bytes, copy_kind(kind)));
  });
}

int vx_memcpy_async(void* dst, const void* src, size_t bytes, int kind, vx_stream_t st) {
  return guard([&] {
    require((dst && src) || bytes == 0, "memcpy: null pointer");
    VX_HIP(hipMemcpyAsync(dst, src, bytes, copy_kind(kind), S(st)));
  });
}

int vx_memset(void* d, int value, size_t bytes) {
  return guard([&] {
    require(d != nullptr || bytes == 0, "memset: null pointer");
    VX_HIP(hipMemset(d, value, bytes));
  });
}

int vx_stream_create(vx_stream_t* out) {
  return guard([&] {
    require(out != nullptr, "stream_create: null out");
    hipStream_t s = nullptr;
    VX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = reinterpret_cast<vx_stream_t>(s);
  });
}

int vx_stream_synchronize(vx_stream_t st) {
  return guard([&] { VX_HIP(hipStreamSynchronize(S(st))); });
}

int vx_stream_destroy(vx_stream_t st) {
  return guard([&] { VX_HIP(hipStreamDestroy(S(st))); });
}

int vx_circle_ntt_evaluate(uint32_t* d, const uint32_t* tw, const uint32_t* ctw, const uint32_t* offs,
                           const uint32_t* sizes, uint32_t nl, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(d != nullptr, "null data");
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    ntt_forward(d, d, n - 1, r, r.log_n, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_circle_ntt_interpolate(uint32_t* d, const uint32_t* tw, const uint32_t* ctw, const uint32_t* offs,
                              const uint32_t* sizes, uint32_t nl, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(d != nullptr, "null data");
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    ntt_inverse(d, r, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

// Column-pointer arrays in the reference's batched calls are device memory
// the caller uploaded first (src/merkle.rs:32-34, src/ntt.rs:226-240). The
// pointers become kernel arguments here, so they are read back to the host:
// after the stream's earlier work (which may have written them), with the
// direction inferred from the address, so a host array is accepted too.
static std::vector<uint32_t*> fetch_ptrs(const void* arr, uint32_t n, hipStream_t s) {
  require(arr != nullptr, "null column-pointer array");
  std::vector<uint32_t*> h(n);
  if (n == 0) return h;
  VX_HIP(hipMemcpyAsync(h.data(), arr, sizeof(void*) * n, hipMemcpyDefault, s));
  VX_HIP(hipStreamSynchronize(s));
  return h;
}

// Columns of a batched in-place transform: the host pointer array is checked
// and copied to the device, in chunks of at most 65535 columns (one grid row
// each), and every pass runs over a whole chunk in one launch. Returns after
// the stream has finished, so the staging buffer outlives the kernels.
static void ntt_cols(uint32_t* const* cols, uint32_t n_cols, const TwiddleRef& r, bool forward, hipStream_t s) {
  if (n_cols == 0) return;
  require(cols != nullptr, "null column array");
  for (uint32_t c = 0; c < n_cols; c++)
    require(cols[c] != nullptr && (reinterpret_cast<uintptr_t>(cols[c]) & 15u) == 0,
            "ntt batch: columns must be non-null and 16-byte aligned");
  constexpr uint32_t CHUNK = 65535;
  DeviceBuf d(sizeof(uint32_t*) * (n_cols < CHUNK ? n_cols : CHUNK));
  for (uint32_t c0 = 0; c0 < n_cols; c0 += CHUNK) {
    const uint32_t m = n_cols - c0 < CHUNK ? n_cols - c0 : CHUNK;
    VX_HIP(hipMemcpyAsync(d.p, cols + c0, sizeof(uint32_t*) * m, hipMemcpyHostToDevice, s));
    uint32_t* const* dc = static_cast<uint32_t* const*>(d.p);
    if (forward)
      ntt_forward_cols(dc, m, r, s);
    else
      ntt_inverse_cols(dc, m, r, s);
    VX_HIP(hipStreamSynchronize(s));
  }
}

int vx_circle_ntt_evaluate_batch(uint32_t* const* d_cols, const uint32_t* tw, const uint32_t* ctw,
                                 const uint32_t* offs, const uint32_t* sizes, uint32_t nl, uint32_t n,
                                 uint32_t n_cols, vx_stream_t st) {
  return guard([&] {
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    if (n_cols == 0) return;
    std::vector<uint32_t*> cols = fetch_ptrs(d_cols, n_cols, S(st));
    ntt_cols(cols.data(), n_cols, r, true, S(st));
  });
}

int vx_circle_ntt_interpolate_batch(uint32_t* const* d_cols, const uint32_t* tw, const uint32_t* ctw,
                                    const uint32_t* offs, const uint32_t* sizes, uint32_t nl, uint32_t n,
                                    uint32_t n_cols, vx_stream_t st) {
  return guard([&] {
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    if (n_cols == 0) return;
    std::vector<uint32_t*> cols = fetch_ptrs(d_cols, n_cols, S(st));
    ntt_cols(cols.data(), n_cols, r, false, S(st));
  });
}

int vx_circle_ntt_evaluate_batch_host(uint32_t* const* h_cols, const uint32_t* tw, const uint32_t* ctw,
                                      const uint32_t* offs, const uint32_t* sizes, uint32_t nl, uint32_t n,
                                      uint32_t n_cols, vx_stream_t st) {
  return guard([&] {
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    ntt_cols(h_cols, n_cols, r, true, S(st));
  });
}

int vx_circle_ntt_interpolate_batch_host(uint32_t* const* h_cols, const uint32_t* tw, const uint32_t* ctw,
                                         const uint32_t* offs, const uint32_t* sizes, uint32_t nl, uint32_t n,
                                         uint32_t n_cols, vx_stream_t st) {
  return guard([&] {
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    ntt_cols(h_cols, n_cols, r, false, S(st));
  });
}

// stwo flat twiddle buffer of an n-point transform (n/2 words, src/ntt.rs:
// 296-354): line layer k at n/2 - 2^(log_n - 1 - k), circle twiddles derived
// from line layer 0 (cuda/circle_ntt_stwo.cu:202-248).
static TwiddleRef stwo_ref(const uint32_t* tw, uint32_t n) {
  require(tw != nullptr, "stwo ntt: null twiddles");
  require(pow2(n) && n >= 4 && n <= (1u << 30), "stwo ntt: n must be a power of two in [4, 2^30]");
  TwiddleRef r{};
  r.line = tw;
  r.log_n = log2u(n);
  r.stwo_circle = 1;
  const uint32_t nl = r.log_n - 1;
  for (uint32_t k = 0; k < nl; k++) r.off[k] = (n >> 1) - (1u << (nl - k));
  return r;
}

int vx_stwo_twiddles(uint32_t log_n, uint32_t* d_tw, uint32_t* d_itw, vx_stream_t st) {
  return guard([&] {
    require(d_tw != nullptr, "stwo_twiddles: null output");
    require(log_n >= 2 && log_n <= 30, "stwo_twiddles: log_n must be 2..30");
    stwo_twiddles(log_n, d_tw, d_itw, S(st));
  });
}

int vx_stwo_ntt_evaluate(uint32_t* d, const uint32_t* tw, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(d != nullptr, "null data");
    TwiddleRef r = stwo_ref(tw, n);
    ntt_forward(d, d, n - 1, r, r.log_n, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_stwo_ntt_interpolate(uint32_t* d, const uint32_t* itw, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(d != nullptr, "null data");
    TwiddleRef r = stwo_ref(itw, n);
    ntt_inverse(d, r, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_stwo_ntt_evaluate_async(uint32_t* d, const uint32_t* tw, uint32_t n, uint32_t n_cols, uint64_t stride,
                               vx_stream_t st) {
  return guard([&] {
    require(d != nullptr && n_cols >= 1 && n_cols <= 65535, "stwo ntt: bad batch");
    TwiddleRef r = stwo_ref(tw, n);
    ntt_forward(d, d, n - 1, r, r.log_n, S(st), n_cols, stride, stride);
  });
}

int vx_stwo_ntt_interpolate_async(uint32_t* d, const uint32_t* itw, uint32_t n, uint32_t n_cols, uint64_t stride,
                                  vx_stream_t st) {
  return guard([&] {
    require(d != nullptr && n_cols >= 1 && n_cols <= 65535, "stwo ntt: bad batch");
    TwiddleRef r = stwo_ref(itw, n);
    ntt_inverse(d, r, S(st), n_cols, stride);
  });
}

int vx_eval_at_point(const uint32_t* d_coeffs, const uint32_t* d_factors, uint32_t* h_result, uint32_t n,
                     uint32_t* d_scratch1, uint32_t* d_scratch2, vx_stream_t st) {
  return guard([&] {
    require(d_coeffs && d_factors && h_result && d_scratch1 && d_scratch2, "eval_at_point: null argument");
    require(pow2(n) && n <= (1u << 31), "eval_at_point: n must be a power of two");
    require(n < 16 || (reinterpret_cast<uintptr_t>(d_coeffs) & 15u) == 0,
            "eval_at_point: coefficients must be 16-byte aligned");
    const uint32_t* r = eval_at_point(d_coeffs, d_factors, log2u(n), d_scratch1, d_scratch2, S(st));
    VX_HIP(hipMemcpyAsync(h_result, r, 16, hipMemcpyDeviceToHost, S(st)));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_eval_at_point_batch(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                           const uint32_t* d_factors, uint32_t max_log_n, uint32_t* d_results, vx_stream_t st) {
  return guard([&] {
    if (n_cols == 0) return;
    require(h_cols && h_log_sizes && d_results, "eval_at_point_batch: null argument");
    require(max_log_n == 0 || d_factors, "eval_at_point_batch: null factors");
    eval_at_point_batch(h_cols, h_log_sizes, n_cols, d_factors, max_log_n, d_results, S(st));
  });
}

int vx_oods_sample(uint32_t* d_channel, const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                   uint32_t trace_log_n, uint32_t* d_points, uint32_t* d_values, vx_stream_t st) {
  return guard([&] {
    require(d_channel && d_points && (n_cols == 0 || (h_cols && h_log_sizes && d_values)),
            "oods_sample: null argument");
    require(trace_log_n <= 31, "oods_sample: trace_log_n must be <= 31");
    for (uint32_t k = 0; k < n_cols; k++)
      require(h_cols[k] && h_log_sizes[k] <= 30, "oods_sample: null column or log size above 30");
    oods_sample(d_channel, h_cols, h_log_sizes, n_cols, trace_log_n, d_points, d_values, S(st));
  });
}

int vx_barycentric_eval(const uint32_t* d_evals, const uint32_t* d_weights, uint32_t n, uint32_t* d_out,
                        uint32_t n_blocks, vx_stream_t st) {
  return guard([&] {
    require(d_out != nullptr && (n == 0 || (d_evals && d_weights)), "barycentric_eval: null argument");
    barycentric_partials(d_evals, d_weights, n, d_out, n_blocks, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_accumulate_numerators(const uint32_t* const* d_col_ptrs, const uint32_t* d_col_indices, const uint32_t* d_b,
                             const uint32_t* d_c, uint32_t n_batch_cols, uint32_t n_rows, uint32_t* o0, uint32_t* o1,
                             uint32_t* o2, uint32_t* o3, vx_stream_t st) {
  return guard([&] {
    if (n_rows == 0) return;
    require(o0 && o1 && o2 && o3, "accumulate_numerators: null output");
    require(n_batch_cols == 0 || (d_col_ptrs && d_col_indices && d_b && d_c), "accumulate_numerators: null input");
    uint32_t* const oz[4] = {o0, o1, o2, o3};
    accumulate_numerators(d_col_ptrs, d_col_indices, nullptr, d_b, d_c, nullptr, nullptr, n_batch_cols, n_rows,
                          false, oz, oz, S(st));
  });
}

int vx_accumulate_numerators_dual(const uint32_t* const* d_col_ptrs, const uint32_t* d_col_indices_z,
                                  const uint32_t* d_zn_coeff_idx, const uint32_t* d_b_z, const uint32_t* d_c_z,
                                  const uint32_t* d_b_zn, const uint32_t* d_c_zn, uint32_t n_cols_z, uint32_t n_rows,
                                  uint32_t* z0, uint32_t* z1, uint32_t* z2, uint32_t* z3, uint32_t* n0, uint32_t* n1,
                                  uint32_t* n2, uint32_t* n3, vx_stream_t st) {
  return guard([&] {
    if (n_rows == 0) return;
    require(z0 && z1 && z2 && z3 && n0 && n1 && n2 && n3, "accumulate_numerators_dual: null output");
    require(n_cols_z == 0 || (d_col_ptrs && d_col_indices_z && d_zn_coeff_idx && d_b_z && d_c_z && d_b_zn && d_c_zn),
            "accumulate_numerators_dual: null input");
    uint32_t* const oz[4] = {z0, z1, z2, z3};
    uint32_t* const on[4] = {n0, n1, n2, n3};
    accumulate_numerators(d_col_ptrs, d_col_indices_z, d_zn_coeff_idx, d_b_z, d_c_z, d_b_zn, d_c_zn, n_cols_z,
                          n_rows, true, oz, on, S(st));
  });
}

int vx_compute_quotients_combine(const uint32_t* d_sp_x, const uint32_t* d_sp_y, const uint32_t* d_fla,
                                 const uint32_t* const* d_numer0, const uint32_t* const* d_numer1,
                                 const uint32_t* const* d_numer2, const uint32_t* const* d_numer3,
                                 const uint32_t* d_acc_log_sizes, uint32_t n_accs, const uint32_t* d_xs,
                                 const uint32_t* d_ys, uint32_t lifting_log_size, uint32_t n_rows, uint32_t* o0,
                                 uint32_t* o1, uint32_t* o2, uint32_t* o3, vx_stream_t st) {
  return guard([&] {
    if (n_rows == 0) return;
    require(o0 && o1 && o2 && o3 && d_xs && d_ys, "compute_quotients_combine: null argument");
    require(n_accs == 0 || (d_sp_x && d_sp_y && d_fla && d_numer0 && d_numer1 && d_numer2 && d_numer3 &&
                            d_acc_log_sizes),
            "compute_quotients_combine: null argument");
    require(lifting_log_size <= 31 && (uint64_t)n_rows <= (1ull << lifting_log_size),
            "compute_quotients_combine: n_rows must not exceed 2^lifting_log_size");
    const uint32_t* const* nu[4] = {d_numer0, d_numer1, d_numer2, d_numer3};
    uint32_t* const out[4] = {o0, o1, o2, o3};
    quotients_combine(d_sp_x, d_sp_y, d_fla, nu, d_acc_log_sizes, n_accs, d_xs, d_ys, lifting_log_size, n_rows, out,
                      S(st));
  });
}

int vx_m31_add(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(n == 0 || (a && b && out), "m31_add: null pointer");
    m31_binop(a, b, out, n, false, S(st));
  });
}

int vx_m31_mul(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(n == 0 || (a && b && out), "m31_mul: null pointer");
    m31_binop(a, b, out, n, true, S(st));
  });
}

int vx_accumulate_lift(uint32_t* col, const uint32_t* curr, uint32_t col_n, uint32_t log_ratio, vx_stream_t st) {
  return guard([&] {
    require(col_n == 0 || (col && curr), "accumulate_lift: null pointer");
    require(log_ratio <= 30, "accumulate_lift: log_ratio must be <= 30");
    accumulate_lift(col, curr, col_n, log_ratio, S(st));
  });
}

int vx_lift_and_accumulate(uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols, vx_stream_t st) {
  return guard([&] {
    require(h_cols && h_log_sizes && n_cols >= 1, "lift_and_accumulate: null argument");
    const uint32_t L = h_log_sizes[n_cols - 1];
    for (uint32_t k = 0; k < n_cols; k++) {
      require(h_log_sizes[k] >= 1 && h_log_sizes[k] <= 31 && (k == 0 || h_log_sizes[k] >= h_log_sizes[k - 1]),
              "lift_and_accumulate: log sizes must be 1..31 and ascending");
      for (uint32_t c = 0; c < 4; c++) require(h_cols[c * n_cols + k] != nullptr, "lift_and_accumulate: null column");
    }
    const uint32_t n_src = n_cols - 1;
    // device tables: src pointers (channel-major), shifts, the 4 last columns
    std::vector<uint64_t> tab(4 * (size_t)n_src + 4 + (n_src + 1) / 2 + 1, 0);
    for (uint32_t c = 0; c < 4; c++)
      for (uint32_t k = 0; k < n_src; k++) tab[c * n_src + k] = reinterpret_cast<uint64_t>(h_cols[c * n_cols + k]);
    for (uint32_t c = 0; c < 4; c++) tab[4 * n_src + c] = reinterpret_cast<uint64_t>(h_cols[c * n_cols + n_src]);
    uint32_t* shifts = reinterpret_cast<uint32_t*>(tab.data() + 4 * n_src + 4);
    for (uint32_t k = 0; k < n_src; k++) shifts[k] = L - h_log_sizes[k] + 1;
    DeviceBuf d(tab.size() * 8);
    VX_HIP(hipMemcpyAsync(d.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, S(st)));
    const uint64_t* dp = static_cast<const uint64_t*>(d.p);
    lift_accumulate(reinterpret_cast<const uint32_t* const*>(dp), reinterpret_cast<const uint32_t*>(dp + 4 * n_src + 4),
                    n_src, reinterpret_cast<uint32_t* const*>(dp + 4 * n_src), 1u << L, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

static int gather_call(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, uint32_t words,
                       bool checked, uint64_t src_len, vx_stream_t st) {
  return guard([&] {
    require(n == 0 || (src && idx && dst), "gather: null pointer");
    gather(src, idx, dst, n, words, checked, src_len, S(st));
  });
}
int vx_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, vx_stream_t st) {
  return gather_call(src, idx, dst, n, 1, false, 0, st);
}
int vx_gather_u32_checked(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, uint32_t src_len,
                          vx_stream_t st) {
  return gather_call(src, idx, dst, n, 1, true, src_len, st);
}
int vx_gather_u256(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, vx_stream_t st) {
  return gather_call(src, idx, dst, n, 8, false, 0, st);
}
int vx_gather_u256_checked(const uint32_t* src, const uint32_t* idx, uint32_t* dst, uint32_t n, uint32_t src_len,
                           vx_stream_t st) {
  return gather_call(src, idx, dst, n, 8, true, src_len, st);
}

static int permute_call(const char* what, const uint32_t* src, uint32_t* dst, uint32_t n, uint32_t log_n,
                        vx_stream_t st) {
  return guard([&] {
    require(log_n <= 31 && (uint64_t)n == (1ull << log_n), (std::string(what) + ": n must be 2^log_n, log_n <= 31").c_str());
    require(src && dst && src != dst, (std::string(what) + ": null or aliasing buffers").c_str());
    permute_hc_brt(src, dst, log_n, S(st));
  });
}
int vx_permute_hc_to_canonic_brt(const uint32_t* src, uint32_t* dst, uint32_t n, uint32_t log_n, vx_stream_t st) {
  return permute_call("permute_hc_to_canonic_brt", src, dst, n, log_n, st);
}
int vx_permute_canonic_brt_to_hc_natural(const uint32_t* src, uint32_t* dst, uint32_t n, uint32_t log_n,
                                         vx_stream_t st) {
  return permute_call("permute_canonic_brt_to_hc_natural", src, dst, n, log_n, st);
}

int vx_pack_leaves(const uint32_t* const* d_in, uint32_t* const* d_out, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(n % 16 == 0, "pack_leaves: N must be a multiple of 16");
    require(n == 0 || (d_in && d_out), "pack_leaves: null pointer array");
    pack_leaves(d_in, d_out, n, S(st));
  });
}

int vx_stwo_ntt_evaluate_lde(uint32_t* d_eval, const uint32_t* d_coeffs, uint32_t coeff_n, const uint32_t* tw,
                             uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(d_eval && d_coeffs && d_eval != d_coeffs, "stwo lde: null or aliasing buffers");
    require(pow2(coeff_n) && coeff_n >= 2 && coeff_n <= n, "stwo lde: coeff_n must be a power of two in [2, n]");
    TwiddleRef r = stwo_ref(tw, n);
    ntt_forward(d_eval, d_coeffs, coeff_n - 1, r, log2u(coeff_n), S(st));
  });
}

int vx_stwo_ntt_evaluate_lde_batch(uint32_t* const* h_eval, const uint32_t* const* h_coeffs, uint32_t coeff_n,
                                   uint32_t n_cols, const uint32_t* tw, uint32_t n, vx_stream_t st) {
  return guard([&] {
    if (n_cols == 0) return;
    require(h_eval && h_coeffs, "stwo lde batch: null column array");
    require(pow2(coeff_n) && coeff_n >= 2 && coeff_n <= n, "stwo lde batch: coeff_n must be a power of two in [2, n]");
    TwiddleRef r = stwo_ref(tw, n);
    for (uint32_t c = 0; c < n_cols; c++) {
      require(h_eval[c] && h_coeffs[c] && h_eval[c] != h_coeffs[c], "stwo lde batch: null or aliasing column");
      require(((reinterpret_cast<uintptr_t>(h_eval[c]) | reinterpret_cast<uintptr_t>(h_coeffs[c])) & 15u) == 0,
              "stwo lde batch: columns must be 16-byte aligned");
    }
    constexpr uint32_t CHUNK = 65535;
    const uint32_t m0 = n_cols < CHUNK ? n_cols : CHUNK;
    DeviceBuf d(2 * sizeof(uint32_t*) * m0);
    for (uint32_t c0 = 0; c0 < n_cols; c0 += CHUNK) {
      const uint32_t m = n_cols - c0 < CHUNK ? n_cols - c0 : CHUNK;
      std::vector<const void*> p(2 * (size_t)m);
      for (uint32_t c = 0; c < m; c++) {
        p[c] = h_eval[c0 + c];
        p[m + c] = h_coeffs[c0 + c];
      }
      VX_HIP(hipMemcpyAsync(d.p, p.data(), p.size() * sizeof(void*), hipMemcpyHostToDevice, S(st)));
      uint32_t* const* dst = static_cast<uint32_t* const*>(d.p);
      ntt_forward_cols(dst, m, r, S(st), reinterpret_cast<const uint32_t* const*>(dst + m), log2u(coeff_n));
      VX_HIP(hipStreamSynchronize(S(st)));
    }
  });
}

int vx_stwo_ntt_batch_evaluate(uint32_t* const* cols, const uint32_t* tw, uint32_t n, uint32_t n_cols,
                               vx_stream_t st) {
  return guard([&] {
    if (n_cols == 0) return;
    TwiddleRef r = stwo_ref(tw, n);
    ntt_cols(cols, n_cols, r, true, S(st));
  });
}

int vx_stwo_ntt_batch_interpolate(uint32_t* const* cols, const uint32_t* itw, uint32_t n, uint32_t n_cols,
                                  vx_stream_t st) {
  return guard([&] {
    if (n_cols == 0) return;
    TwiddleRef r = stwo_ref(itw, n);
    ntt_cols(cols, n_cols, r, false, S(st));
  });
}

int vx_circle_ntt_layer(uint32_t* d, const uint32_t* tw, uint32_t li, uint32_t n, int fwd, vx_stream_t st) {
  return guard([&] {
    require(d && tw && pow2(n) && n >= 2 && li < log2u(n), "ntt_layer: bad arguments");
    ntt_layer(d, tw, li, n, fwd != 0, S(st));
  });
}

int vx_circle_ntt_evaluate_lde(uint32_t* d_eval, const uint32_t* d_coeffs, uint32_t coeff_n, const uint32_t* tw,
                               const uint32_t* ctw, const uint32_t* offs, uint32_t nl, uint32_t n,
                               vx_stream_t st) {
  return guard([&] {
    require(d_eval && d_coeffs && pow2(coeff_n) && coeff_n >= 2 && coeff_n <= n, "lde: bad sizes");
    require(d_eval != d_coeffs, "lde: output must not alias the coefficients");
    TwiddleRef r = tw_ref(tw, ctw, offs, nl, n);
    ntt_forward(d_eval, d_coeffs, coeff_n - 1, r, log2u(coeff_n), S(st));
  });
}

int vx_bit_reverse_m31(uint32_t* d, uint32_t log_n, vx_stream_t st) {
  return guard([&] {
    require(d && log_n <= 31, "bit_reverse: bad arguments");
    bit_reverse_m31(d, log_n, S(st));
  });
}

int vx_bit_reverse_qm31(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t log_n, vx_stream_t st) {
  return guard([&] {
    require(in && out && log_n <= 31 && n == (1u << log_n), "bit_reverse_qm31: n must be 2^log_n");
    bit_reverse_qm31(in, out, log_n, S(st));
  });
}

int vx_compute_coset_points(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* ox, uint32_t* oy,
                            uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(ox || oy, "coset_points: no output");
    coset_points({ix, iy}, {sx, sy}, ox, oy, n, S(st));
  });
}

int vx_extract_and_squash(const uint32_t* in, uint32_t* tw, uint32_t* sq, uint32_t half_n, vx_stream_t st) {
  return guard([&] {
    require(in && tw && sq, "extract_and_squash: null pointer");
    extract_and_squash(in, tw, sq, half_n, S(st));
  });
}

int vx_squash_x(const uint32_t* in, uint32_t* out, uint32_t out_n, vx_stream_t st) {
  return guard([&] {
    require(in && out, "squash_x: null pointer");
    extract_and_squash(in, nullptr, out, out_n, S(st));
  });
}

int vx_batch_inverse_m31(const uint32_t* in, uint32_t* out, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(in && out, "batch_inverse: null pointer");
    batch_inverse(in, out, n, S(st));
  });
}

int vx_compute_fold_twiddle_sources(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* out, uint32_t n,
                                    uint32_t log_n, int extract_y, vx_stream_t st) {
  return guard([&] {
    require(out != nullptr, "fold_twiddle_sources: null output");
    fold_twiddle_sources({ix, iy}, {sx, sy}, out, n, log_n, extract_y != 0, S(st));
  });
}

int vx_fibonacci_trace(uint32_t a, uint32_t b, uint32_t* out, uint32_t log_n, vx_stream_t st) {
  return guard([&] {
    require(out != nullptr, "fibonacci_trace: null output");
    fib_trace(a, b, out, log_n, S(st));
  });
}

int vx_fibonacci_quotient(const uint32_t* t, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3,
                          const uint32_t* alpha, uint32_t n, vx_stream_t st) {
  return guard([&] {
    require(t && o0 && o1 && o2 && o3, "quotient: null pointer");
    fib_quotient(t, o0, o1, o2, o3, host_q(alpha), 0, n, n, S(st));
  });
}

int vx_fibonacci_quotient_chunk(const uint32_t* t, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3,
                                const uint32_t* alpha, uint32_t offset, uint32_t chunk_n, uint32_t global_n,
                                vx_stream_t st) {
  return guard([&] {
    require(t && o0 && o1 && o2 && o3, "quotient_chunk: null pointer");
    fib_quotient(t, o0, o1, o2, o3, host_q(alpha), offset, chunk_n, global_n, S(st));
  });
}

int vx_zero_pad(const uint32_t* src, uint32_t* dst, uint32_t src_n, uint32_t dst_n, vx_stream_t st) {
  return guard([&] {
    require(dst && (src || src_n == 0), "zero_pad: null pointer");
    zero_pad(src, dst, src_n, dst_n, S(st));
  });
}

int vx_fold_line_soa(const uint32_t* i0, const uint32_t* i1, const uint32_t* i2, const uint32_t* i3,
                     const uint32_t* tw, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3,
                     const uint32_t* alpha, uint32_t half_n, vx_stream_t st) {
  return guard([&] {
    require(i0 && i1 && i2 && i3 && tw && o0 && o1 && o2 && o3, "fold_line: null pointer");
    const uint32_t* in[4] = {i0, i1, i2, i3};
    uint32_t* out[4] = {o0, o1, o2, o3};
    fold_line(in, tw, out, host_q(alpha), half_n, S(st));
  });
}

int vx_fold_circle_into_line_soa(uint32_t* d0, uint32_t* d1, uint32_t* d2, uint32_t* d3, const uint32_t* s0,
                                 const uint32_t* s1, const uint32_t* s2, const uint32_t* s3, const uint32_t* tw,
                                 const uint32_t* alpha, const uint32_t* alpha_sq, uint32_t half_n,
                                 vx_stream_t st) {
  return guard([&] {
    require(d0 && d1 && d2 && d3 && s0 && s1 && s2 && s3 && tw, "fold_circle: null pointer");
    uint32_t* dst[4] = {d0, d1, d2, d3};
    const uint32_t* src[4] = {s0, s1, s2, s3};
    fold_circle_into_line(dst, src, tw, host_q(alpha), host_q(alpha_sq), half_n, S(st));
  });
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

static LeafSrc cols_leaf(const uint32_t* const* cols, uint32_t n_cols) {
  require(cols != nullptr && n_cols >= 1, "merkle: need at least one column");
  LeafSrc L{};
  L.kind = LEAF_COLS;
  L.n_cols = n_cols > 16 ? 16 : n_cols;
  for (uint32_t c = 0; c < L.n_cols; c++) {
    require(cols[c] != nullptr, "merkle: null column");
    L.col[c] = cols[c];
  }
  // t = 4 * min(n_cols, 16): extra columns are ignored (cuda/blake2s.cu:50-54)
  return L;
}

int vx_merkle_hash_leaves_host(const uint32_t* const* cols, uint32_t* hashes, uint32_t n_cols, uint32_t n_leaves,
                               vx_stream_t st) {
  return guard([&] {
    require(hashes != nullptr, "hash_leaves: null output");
    merkle_leaf_hashes(cols_leaf(cols, n_cols), n_leaves, hashes, S(st));
  });
}

int vx_merkle_hash_leaves(const uint32_t* const* d_cols, uint32_t* hashes, uint32_t n_cols, uint32_t n_leaves,
                          vx_stream_t st) {
  return guard([&] {
    require(hashes != nullptr && n_cols >= 1, "hash_leaves: null output or no column");
    // only the first 16 columns enter a leaf (cuda/blake2s.cu:50-54)
    std::vector<uint32_t*> cols = fetch_ptrs(d_cols, n_cols > 16 ? 16 : n_cols, S(st));
    merkle_leaf_hashes(cols_leaf(cols.data(), n_cols), n_leaves, hashes, S(st));
  });
}

int vx_merkle_hash_leaves_single(const uint32_t* col, uint32_t* hashes, uint32_t n_leaves, vx_stream_t st) {
  const uint32_t* cols[1] = {col};
  return vx_merkle_hash_leaves_host(cols, hashes, 1, n_leaves, st);
}

int vx_merkle_hash_leaves_quad(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                               uint32_t* hashes, uint32_t n_leaves, vx_stream_t st) {
  const uint32_t* cols[4] = {c0, c1, c2, c3};
  return vx_merkle_hash_leaves_host(cols, hashes, 4, n_leaves, st);
}

int vx_merkle_hash_nodes(const uint32_t* ch, uint32_t* par, uint32_t n_parents, vx_stream_t st) {
  return guard([&] {
    require(ch && par, "hash_nodes: null pointer");
    merkle_hash_nodes(ch, par, n_parents, S(st));
  });
}

int vx_merkle_reduce_to_root(const uint32_t* nodes, uint32_t* root, uint32_t n_nodes, vx_stream_t st) {
  return guard([&] {
    require(nodes && root, "reduce_to_root: null pointer");
    merkle_reduce_to_root(nodes, root, n_nodes, S(st));
  });
}

int vx_merkle_commit_small_soa4(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                                uint32_t* root, uint32_t n_leaves, vx_stream_t st) {
  return guard([&] {
    require(root && pow2(n_leaves) && n_leaves >= 4 && n_leaves <= 2048, "commit_small: 4..2048 leaves");
    const uint32_t* cols[4] = {c0, c1, c2, c3};
    LeafSrc L = cols_leaf(cols, 4);
    for (int c = 0; c < 4; c++) require(aligned16(L.col[c]), "commit_small: columns must be 16-byte aligned");
    merkle_root_small(L, log2u(n_leaves), root, S(st));
  });
}

int vx_merkle_hash_leaves_merge_soa4(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                                     uint32_t* parents, uint32_t n_leaves, vx_stream_t st) {
  return guard([&] {
    require(parents != nullptr, "leaves_merge: null output");
    const uint32_t* cols[4] = {c0, c1, c2, c3};
    merkle_leaf_pairs(cols_leaf(cols, 4), n_leaves, parents, S(st));
  });
}

// Root of every 1024-leaf tile of n_leaves (a multiple of 1024): the tiles
// are covered by one subtree launch per set bit of n_leaves / 1024, largest
// first, each over its own row range (tile roots stay in tile order).
static void tiled_roots(const LeafSrc& L0, uint32_t n_leaves, uint32_t* roots, hipStream_t s) {
  require(roots != nullptr && n_leaves >= 1024 && n_leaves % 1024 == 0,
          "tiled: n_leaves must be a positive multiple of 1024");
  for (uint32_t c = 0; c < L0.n_cols; c++) require(aligned16(L0.col[c]), "tiled: columns must be 16-byte aligned");
  const uint32_t tiles = n_leaves / 1024;
  uint64_t done = 0;
  for (int b = 31; b >= 0; b--) {
    if (!((tiles >> b) & 1u)) continue;
    LeafSrc L = L0;
    for (uint32_t c = 0; c < L.n_cols; c++) L.col[c] = L0.col[c] + done * 1024;
    merkle_tiled(L, 10 + (uint32_t)b, 10, roots + 8 * done, s);
    done += 1ull << b;
  }
}

int vx_merkle_tiled_generic_host(const uint32_t* const* cols, uint32_t* roots, uint32_t n_cols, uint32_t n_leaves,
                                 vx_stream_t st) {
  return guard([&] { tiled_roots(cols_leaf(cols, n_cols), n_leaves, roots, S(st)); });
}

int vx_merkle_tiled_generic(const uint32_t* const* d_cols, uint32_t* roots, uint32_t n_cols, uint32_t n_leaves,
                            vx_stream_t st) {
  return guard([&] {
    require(n_cols >= 1, "tiled: need at least one column");
    std::vector<uint32_t*> cols = fetch_ptrs(d_cols, n_cols > 16 ? 16 : n_cols, S(st));
    tiled_roots(cols_leaf(cols.data(), n_cols), n_leaves, roots, S(st));
  });
}

int vx_merkle_tiled_soa4(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                         uint32_t* roots, uint32_t n_leaves, vx_stream_t st) {
  const uint32_t* cols[4] = {c0, c1, c2, c3};
  return vx_merkle_tiled_generic_host(cols, roots, 4, n_leaves, st);
}

int vx_grind_pow(const uint32_t* prefix, uint64_t* result, uint32_t pow_bits, uint64_t batch_offset,
                 uint32_t n_threads, vx_stream_t st) {
  return guard([&] {
    require(prefix && result && n_threads > 0, "grind_pow: bad arguments");
    grind_pow_batch(prefix, result, pow_bits, batch_offset, n_threads, false, S(st));
  });
}

int vx_grind_pow_m31_output(const uint32_t* prefix, uint64_t* result, uint32_t pow_bits, uint64_t batch_offset,
                            uint32_t n_threads, vx_stream_t st) {
  return guard([&] {
    require(prefix && result && n_threads > 0, "grind_pow_m31_output: bad arguments");
    grind_pow_batch(prefix, result, pow_bits, batch_offset, n_threads, true, S(st));
  });
}

static void grind_stwo_call(const uint32_t* prefix, uint32_t pow_bits, uint64_t* result, bool m31_out,
                            hipStream_t s) {
  require(prefix && result, "grind_pow_stwo: null pointer");
  if (pow_bits == 0) {
    const uint64_t zero = 0;
    VX_HIP(hipMemcpyAsync(result, &zero, 8, hipMemcpyHostToDevice, s));
    VX_HIP(hipStreamSynchronize(s));
    return;
  }
  grind_pow_stwo(prefix, pow_bits, result, m31_out, s);
}

int vx_grind_pow_stwo(const uint32_t* prefix, uint32_t pow_bits, uint64_t* result, vx_stream_t st) {
  return guard([&] { grind_stwo_call(prefix, pow_bits, result, false, S(st)); });
}

int vx_grind_pow_stwo_m31(const uint32_t* prefix, uint32_t pow_bits, uint64_t* result, vx_stream_t st) {
  return guard([&] { grind_stwo_call(prefix, pow_bits, result, true, S(st)); });
}

int vx_grind_pow_poseidon(const uint64_t* prefix_mont, uint64_t* result, uint32_t pow_bits, uint64_t batch_offset,
                          uint32_t n_threads, vx_stream_t st) {
  return guard([&] { poseidon_grind_batch(prefix_mont, result, pow_bits, batch_offset, n_threads, S(st)); });
}

int vx_grind_pow_stwo_poseidon(const uint64_t* prefix_mont, uint32_t pow_bits, uint64_t* result, vx_stream_t st) {
  return guard([&] {
    require(prefix_mont && result, "grind_pow_stwo_poseidon: null pointer");
    if (pow_bits == 0) {
      const uint64_t zero = 0;
      VX_HIP(hipMemcpyAsync(result, &zero, 8, hipMemcpyHostToDevice, S(st)));
      VX_HIP(hipStreamSynchronize(S(st)));
      return;
    }
    poseidon_grind_min(prefix_mont, pow_bits, result, S(st));
  });
}

int vx_build_leaves_poseidon252(const uint32_t* const* d_col_ptrs, const uint32_t* d_col_log_sizes, uint32_t n_cols,
                                uint32_t lifting_log_size, uint64_t* d_out, uint32_t n_leaves, vx_stream_t st) {
  return guard([&] {
    poseidon_leaves(d_col_ptrs, d_col_log_sizes, n_cols, lifting_log_size, d_out, n_leaves, S(st));
  });
}

int vx_build_next_layer_poseidon252(const uint64_t* d_prev, uint64_t* d_out, uint32_t n_parents, vx_stream_t st) {
  return guard([&] { poseidon_next_layer(d_prev, d_out, n_parents, S(st)); });
}

int vx_grind_pow_min(const uint32_t* prefix, uint32_t pow_bits, uint64_t* result, vx_stream_t st) {
  return guard([&] {
    require(prefix && result, "grind_pow_min: null pointer");
    grind_pow(prefix, pow_bits, result, S(st));
  });
}

// ----------------------------------------------------------- prove / verify
struct vx_prover {
  StarkProver impl;
  vx_prover(uint32_t l, uint32_t b, uint32_t q, uint32_t p, hipStream_t s, const AirSpec& air = {})
      : impl(l, b, q, p, s, air) {}
};

size_t vx_proof_size(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits) {
  if (log_n < 1 || blowup_bits < 1 || log_n + blowup_bits > 31 || log_n + blowup_bits < 4) return 0;
  return proof_size(log_n, blowup_bits, nq, pow_bits);
}

int vx_prover_create(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits, vx_stream_t st,
                     vx_prover** out) {
  return guard([&] {
    require(out != nullptr, "prover_create: null out");
    *out = nullptr;
    *out = new vx_prover(log_n, blowup_bits, nq, pow_bits, S(st));
  });
}

int vx_d2_prover_create(uint32_t log_n, uint32_t n_cols, uint32_t n_cons, uint64_t air_seed, uint32_t blowup_bits,
                        uint32_t nq, uint32_t pow_bits, vx_stream_t st, vx_prover** out) {
  return guard([&] {
    require(out != nullptr, "prover_create: null out");
    *out = nullptr;
    AirSpec air;
    air.kind = AIR_D2;
    air.n_cols = n_cols;
    air.n_cons = n_cons;
    air.air_seed = air_seed;
    *out = new vx_prover(log_n, blowup_bits, nq, pow_bits, S(st), air);
  });
}

size_t vx_d2_proof_size(uint32_t log_n, uint32_t n_cols, uint32_t blowup_bits, uint32_t nq) {
  if (log_n < 1 || blowup_bits < 1 || log_n + blowup_bits > 31 || log_n + blowup_bits < 4 || n_cols < 1 ||
      n_cols > D2_MAX_COLS)
    return 0;
  AirSpec air;
  air.kind = AIR_D2;
  air.n_cols = n_cols;
  return proof_size(log_n, blowup_bits, nq, 1, air);
}

int vx_batch_prover_create(const uint32_t* h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t nq,
                           uint32_t pow_bits, vx_stream_t st, vx_prover** out) {
  return guard([&] {
    require(out != nullptr && h_log_sizes != nullptr && n_cols >= 1, "batch_prover_create: bad arguments");
    *out = nullptr;
    AirSpec air;
    air.kind = AIR_BATCH;
    air.n_cols = n_cols;
    air.log_sizes.assign(h_log_sizes, h_log_sizes + n_cols);
    *out = new vx_prover(h_log_sizes[n_cols - 1], blowup_bits, nq, pow_bits, S(st), air);
  });
}

size_t vx_batch_proof_size(const uint32_t* h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t nq) {
  if (!h_log_sizes || n_cols < 1 || n_cols > 4096 || blowup_bits < 1) return 0;
  const uint32_t S = h_log_sizes[n_cols - 1] + blowup_bits;
  if (S < 4 || S > 31) return 0;
  AirSpec air;
  air.kind = AIR_BATCH;
  air.n_cols = n_cols;
  air.log_sizes.assign(h_log_sizes, h_log_sizes + n_cols);
  return proof_size(h_log_sizes[n_cols - 1], blowup_bits, nq, 1, air);
}

int vx_prover_destroy(vx_prover* p) {
  return guard([&] { delete p; });
}

int vx_prover_enqueue(vx_prover* p, uint32_t a, uint32_t b, int timed) {
  return guard([&] {
    require(p != nullptr, "null prover");
    p->impl.enqueue(a, b, timed != 0);
  });
}

int vx_prover_wait(vx_prover* p, uint8_t* out, size_t cap, size_t* len) {
  return guard([&] {
    require(p != nullptr, "null prover");
    size_t n = 0;
    const uint8_t* src = p->impl.wait(&n);
    if (len) *len = n;
    if (out) {
      require(cap >= n, "output buffer smaller than the proof (see vx_proof_size)");
      std::memcpy(out, src, n);
    }
  });
}

int vx_prover_prove(vx_prover* p, uint32_t a, uint32_t b, uint8_t* out, size_t cap, size_t* len) {
  int rc = vx_prover_enqueue(p, a, b, 0);
  if (rc) return rc;
  return vx_prover_wait(p, out, cap, len);
}

int vx_prover_stage_times(vx_prover* p, float out[11]) {
  return guard([&] {
    require(p && out, "null argument");
    StageTimes t = p->impl.last_times();
    float v[11] = {t.trace_ms, t.interpolate_ms, t.lde_ms,     t.trace_commit_ms, t.quotient_commit_ms, t.fri_ms,
                   t.queries_ms, t.decommit_ms,  t.download_ms, t.total_ms,       t.trace_subtree_ms};
    std::memcpy(out, v, sizeof v);
  });
}

uint64_t vx_prover_device_bytes(const vx_prover* p) { return p ? p->impl.device_bytes() : 0; }

int vx_prover_buffer(const vx_prover* p, int which, const uint32_t** ptr, uint64_t* words) {
  return guard([&] {
    require(p && ptr && words, "null argument");
    const ProofLayout& L = p->impl.layout();
    if (which == 0) {
      *ptr = p->impl.d_trace();
      *words = 1ull << L.log_n;
    } else if (which == 1) {
      *ptr = p->impl.d_eval();
      *words = 1ull << L.le;
    } else if (which == 100) {
      *ptr = p->impl.d_proof();
      *words = L.total_words;
    } else if (which == 101) {
      *ptr = reinterpret_cast<const uint32_t*>(p->impl.host_proof());
      *words = L.total_words;
    } else if (which >= 2 && (uint32_t)(which - 2) < L.n_fri) {
      *ptr = p->impl.d_layer(which - 2);
      *words = 4ull << (L.le - 1 - (which - 2));
    } else {
      require(false, "prover_buffer: unknown buffer");
    }
  });
}

int vx_prove(uint32_t a, uint32_t b, uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits,
             uint8_t* out, size_t cap, size_t* len) {
  vx_prover* p = nullptr;
  int rc = vx_prover_create(log_n, blowup_bits, nq, pow_bits, nullptr, &p);
  if (rc) return rc;
  rc = vx_prover_prove(p, a, b, out, cap, len);
  vx_prover_destroy(p);
  return rc;
}

int vx_verify(const uint8_t* proof, size_t len, uint32_t blowup_bits, uint32_t nq, char* err, size_t err_cap) {
  int rc = guard([&] {
    require(proof != nullptr || len == 0, "verify: null proof");
    std::string why;
    if (!verify_proof(proof, len, blowup_bits, nq, why)) throw Error(ERR_VERIFY, why);
  });
  if (err && err_cap) {
    std::string m = rc == VX_OK ? "" : g_last_error;
    std::strncpy(err, m.c_str(), err_cap - 1);
    err[err_cap - 1] = 0;
  }
  return rc;
}

// ------------------------------------------------------- stwo export
int vx_export_stwo_json(const uint8_t* proof, size_t len, uint32_t blowup_bits, char* out, size_t cap,
                        size_t* out_len) {
  return guard([&] {
    const std::string j = export_stwo_json(proof, len, blowup_bits);
    if (out_len) *out_len = j.size();
    if (out) {
      require(cap > j.size(), "stwo export: output buffer too small (query the size with out = NULL)");
      std::memcpy(out, j.data(), j.size());
      out[j.size()] = 0;
    }
  });
}

static std::vector<uint64_t> positions_of(const uint64_t* p, size_t n) {
  require(n == 0 || p != nullptr, "null positions");
  return std::vector<uint64_t>(p, p + n);
}

int vx_stwo_hash_witness(const uint64_t* positions, size_t n, const uint32_t* paths, uint32_t height,
                         uint32_t* out, size_t cap, size_t* out_n) {
  return guard([&] {
    require(height <= 64 && (n == 0 || height == 0 || paths != nullptr), "hash witness: bad paths");
    std::vector<const uint32_t*> ps(n);
    std::vector<size_t> words(n, 8ull * height);
    for (size_t i = 0; i < n; i++) ps[i] = paths + 8ull * height * i;
    const auto w = hash_witness(positions_of(positions, n), ps, words, height);
    if (out_n) *out_n = w.size();
    if (out) {
      require(cap >= w.size(), "hash witness: output smaller than the witness");
      for (size_t i = 0; i < w.size(); i++) std::memcpy(out + 8 * i, w[i].data(), 32);
    }
  });
}

int vx_stwo_fri_witness(const uint64_t* positions, size_t n, const uint32_t* sibling_values, uint32_t* out,
                        size_t cap, size_t* out_n) {
  return guard([&] {
    require(n == 0 || sibling_values != nullptr, "fri witness: null sibling values");
    const auto w = fri_witness(positions_of(positions, n), sibling_values);
    if (out_n) *out_n = w.size();
    if (out) {
      require(cap >= w.size(), "fri witness: output smaller than the witness");
      for (size_t i = 0; i < w.size(); i++) std::memcpy(out + 4 * i, w[i].data(), 16);
    }
  });
}

int vx_stwo_reconstruct_root(uint32_t log_height, const uint64_t* positions, size_t n, const uint32_t* col_values,
                             uint32_t n_cols, const uint32_t* witness, size_t n_witness, uint32_t root[8], char* err,
                             size_t err_cap) {
  int rc = guard([&] {
    require(log_height <= 64 && n_cols >= 1 && n_cols <= (1u << 16), "reconstruct root: bad shape");
    require((n == 0 || col_values != nullptr) && (n_witness == 0 || witness != nullptr) && root != nullptr,
            "reconstruct root: null argument");
    const std::string why =
        reconstruct_root(log_height, positions_of(positions, n), col_values, n_cols, witness, n_witness, root);
    if (!why.empty()) throw Error(ERR_VERIFY, why);
  });
  if (err && err_cap) {
    std::string m = rc == VX_OK ? "" : g_last_error;
    std::strncpy(err, m.c_str(), err_cap - 1);
    err[err_cap - 1] = 0;
  }
  return rc;
}

int vx_stwo_last_layer_coeffs(const uint32_t* evals, uint32_t n, uint32_t* out) {
  return guard([&] {
    require(evals != nullptr && out != nullptr, "last layer: null argument");
    const auto c = last_layer_coeffs(evals, n);
    for (uint32_t i = 0; i < n; i++) std::memcpy(out + 4 * i, c[i].data(), 16);
  });
}

int vx_prover_set_graphs(vx_prover* p, int on) {
  return guard([&] {
    require(p != nullptr, "null prover");
    p->impl.set_graphs(on != 0);
  });
}

// ------------------------------------------------------- fused commits
// Tree levels store_lo..log_n back to back from d_levels (8 words per node).
static TreeLevels caller_levels(uint32_t* d_levels, uint32_t log_n, uint32_t store_lo) {
  require(d_levels != nullptr && log_n >= 2 && log_n <= 31 && store_lo >= 2 && store_lo <= log_n,
          "commit: bad tree shape (need 2 <= store_lo <= log_n <= 31)");
  TreeLevels T{};
  T.base = d_levels;
  T.log_n = log_n;
  T.lo = store_lo;
  uint64_t w = 0;
  for (uint32_t l = store_lo; l <= log_n; l++) {
    T.off[l] = w;
    w += 8ull << (log_n - l);
  }
  return T;
}

uint64_t vx_merkle_levels_words(uint32_t log_n, uint32_t store_lo) {
  if (log_n > 31 || store_lo > log_n) return 0;
  uint64_t w = 0;
  for (uint32_t l = store_lo; l <= log_n; l++) w += 8ull << (log_n - l);
  return w;
}

// Alphas arrive as host [4] words (the reference passes them by value); the
// fused producers read them from device memory, so they are staged in a
// small buffer and the call waits for its stream before returning.
struct StagedAlphas {
  DeviceBuf buf{64};
  uint32_t* put(int k, const uint32_t* a, hipStream_t s) {
    require(a != nullptr, "commit: null alpha");
    uint32_t* d = buf.u32() + 4 * k;
    VX_HIP(hipMemcpyAsync(d, a, 16, hipMemcpyHostToDevice, s));
    return d;
  }
};

int vx_merkle_commit_tiled(const uint32_t* const* h_cols, uint32_t n_cols, uint32_t log_n, uint32_t store_lo,
                           uint32_t* d_levels, uint32_t* d_root, vx_stream_t st) {
  return guard([&] {
    require(d_root != nullptr && n_cols >= 1 && n_cols <= 16, "commit_tiled: 1..16 columns and a root");
    const TreeLevels T = caller_levels(d_levels, log_n, store_lo);
    RootHook hook{};
    hook.root_out = d_root;
    merkle_commit(cols_leaf(h_cols, n_cols), T, store_lo, hook, S(st));
  });
}

int vx_quotient_commit(const uint32_t* d_eval, const uint32_t* alpha, uint32_t log_n, uint32_t store_lo,
                       uint32_t* d_levels, uint32_t* d_root, vx_stream_t st) {
  return guard([&] {
    require(d_eval != nullptr && d_root != nullptr, "quotient_commit: null pointer");
    const TreeLevels T = caller_levels(d_levels, log_n, store_lo);
    StagedAlphas a;
    LeafSrc L{};
    L.kind = LEAF_QUOTIENT;
    L.col[0] = d_eval;
    L.alpha_q = a.put(0, alpha, S(st));
    L.src_mask = (1ull << log_n) - 1;
    RootHook hook{};
    hook.root_out = d_root;
    merkle_commit(L, T, store_lo, hook, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_circle_fold_commit(const uint32_t* d_eval, const uint32_t* alpha_q, const uint32_t* alpha_f,
                          const uint32_t* d_twiddles, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3,
                          uint32_t log_n, uint32_t store_lo, uint32_t* d_levels, uint32_t* d_root, vx_stream_t st) {
  return guard([&] {
    require(d_eval && d_twiddles && o0 && o1 && o2 && o3 && d_root && log_n >= 3,
            "circle_fold_commit: null pointer or log_n < 3");
    const TreeLevels T = caller_levels(d_levels, log_n - 1, store_lo);
    StagedAlphas a;
    LeafSrc L{};
    L.kind = LEAF_FOLD_CIRCLE_Q;
    L.col[0] = d_eval;
    L.alpha_q = a.put(0, alpha_q, S(st));
    L.alpha_f = a.put(1, alpha_f, S(st));
    L.tw = d_twiddles;
    L.src_mask = (1ull << log_n) - 1;
    uint32_t* out[4] = {o0, o1, o2, o3};
    for (int c = 0; c < 4; c++) {
      require(aligned16(out[c]), "circle_fold_commit: outputs must be 16-byte aligned");
      L.out[c] = out[c];
    }
    RootHook hook{};
    hook.root_out = d_root;
    merkle_commit(L, T, store_lo, hook, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_fold_line_commit(const uint32_t* i0, const uint32_t* i1, const uint32_t* i2, const uint32_t* i3,
                        const uint32_t* d_twiddles, uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3,
                        const uint32_t* alpha, uint32_t log_half_n, uint32_t store_lo, uint32_t* d_levels,
                        uint32_t* d_root, vx_stream_t st) {
  return guard([&] {
    require(i0 && i1 && i2 && i3 && d_twiddles && o0 && o1 && o2 && o3 && d_root, "fold_line_commit: null pointer");
    const TreeLevels T = caller_levels(d_levels, log_half_n, store_lo);
    StagedAlphas a;
    LeafSrc L{};
    L.kind = LEAF_FOLD_LINE;
    const uint32_t* in[4] = {i0, i1, i2, i3};
    uint32_t* out[4] = {o0, o1, o2, o3};
    for (int c = 0; c < 4; c++) {
      require(aligned16(in[c]) && aligned16(out[c]), "fold_line_commit: columns must be 16-byte aligned");
      L.col[c] = in[c];
      L.out[c] = out[c];
    }
    L.tw = d_twiddles;
    require(aligned16(d_twiddles), "fold_line_commit: twiddles must be 16-byte aligned");
    L.alpha_f = a.put(0, alpha, S(st));
    RootHook hook{};
    hook.root_out = d_root;
    merkle_commit(L, T, store_lo, hook, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

// ------------------------------------------------------------ autotuner
int vx_tuner_configure(int enabled, const char* cache_file) {
  return guard([&] { tuner_configure(enabled != 0, cache_file); });
}

int vx_tuner_entries(uint32_t* out, uint32_t capacity, uint32_t* n) {
  return guard([&] {
    require(n != nullptr && (out != nullptr || capacity == 0), "tuner: null argument");
    const std::vector<TunedEntry> e = tuner_entries();
    *n = (uint32_t)e.size();
    for (uint32_t i = 0; i < e.size() && i < capacity; i++) {
      const TunedEntry& t = e[i];
      const uint32_t w[9] = {t.device, t.kind, t.n_cols, t.n_cons, t.lift_n, t.log_n, t.store_lo, t.tile_log, t.lpt};
      for (int k = 0; k < 9; k++) out[9 * i + k] = w[k];
    }
  });
}

int vx_tuner_ntt_entries(uint32_t* out, uint32_t capacity, uint32_t* n) {
  return guard([&] {
    require(n != nullptr && (out != nullptr || capacity == 0), "tuner: null argument");
    const std::vector<uint32_t> e = tuner_ntt_entries();
    *n = (uint32_t)(e.size() / 8);
    for (size_t i = 0; i < e.size() && i < 8ull * capacity; i++) out[i] = e[i];
  });
}

// ------------------------------------------------------------------ GKR
int vx_gkr_fix_first_variable_base(const uint32_t* in, uint32_t* out, const uint32_t* r, uint32_t n, vx_stream_t st) {
  return guard([&] { gkr_fix_first_variable(in, out, r, n, false, S(st)); });
}
int vx_gkr_fix_first_variable_secure(const uint32_t* in, uint32_t* out, const uint32_t* r, uint32_t n,
                                     vx_stream_t st) {
  return guard([&] { gkr_fix_first_variable(in, out, r, n, true, S(st)); });
}
int vx_gkr_gen_eq_evals_init(uint32_t* buf, const uint32_t* v, vx_stream_t st) {
  return guard([&] { gkr_gen_eq_evals_init(buf, v, S(st)); });
}
int vx_gkr_gen_eq_evals_pass(uint32_t* buf, const uint32_t* y, uint32_t cur, vx_stream_t st) {
  return guard([&] { gkr_gen_eq_evals_pass(buf, y, cur, S(st)); });
}
int vx_gkr_gen_eq_evals(uint32_t* buf, const uint32_t* y, uint32_t k, const uint32_t* v, vx_stream_t st) {
  return guard([&] { gkr_gen_eq_evals(buf, y, k, v, S(st)); });
}
int vx_gkr_next_layer_grand_product(const uint32_t* in, uint32_t* out, uint32_t n, vx_stream_t st) {
  return guard([&] { gkr_next_layer(0, in, nullptr, out, nullptr, n, S(st)); });
}
int vx_gkr_next_layer_logup_generic(const uint32_t* num, const uint32_t* den, uint32_t* on, uint32_t* od, uint32_t n,
                                    vx_stream_t st) {
  return guard([&] { gkr_next_layer(1, num, den, on, od, n, S(st)); });
}
int vx_gkr_next_layer_logup_mult(const uint32_t* num, const uint32_t* den, uint32_t* on, uint32_t* od, uint32_t n,
                                 vx_stream_t st) {
  return guard([&] { gkr_next_layer(2, num, den, on, od, n, S(st)); });
}
int vx_gkr_next_layer_logup_singles(const uint32_t* den, uint32_t* on, uint32_t* od, uint32_t n, vx_stream_t st) {
  return guard([&] { gkr_next_layer(3, nullptr, den, on, od, n, S(st)); });
}
static int gkr_sum_call(int kind, const uint32_t* eq, const uint32_t* num, const uint32_t* den,
                        const uint32_t* lambda, uint32_t* partial, uint32_t n_terms, uint32_t* n_blocks,
                        vx_stream_t st) {
  return guard([&] {
    require(n_blocks != nullptr, "gkr_sum_poly: null n_blocks");
    *n_blocks = gkr_sum_poly(kind, eq, num, den, lambda, partial, n_terms, S(st));
  });
}
int vx_gkr_sum_poly_grand_product(const uint32_t* eq, const uint32_t* layer, uint32_t* partial, uint32_t n_terms,
                                  uint32_t* n_blocks, vx_stream_t st) {
  return gkr_sum_call(0, eq, layer, nullptr, nullptr, partial, n_terms, n_blocks, st);
}
int vx_gkr_sum_poly_logup_generic(const uint32_t* eq, const uint32_t* num, const uint32_t* den, const uint32_t* lambda,
                                  uint32_t* partial, uint32_t n_terms, uint32_t* n_blocks, vx_stream_t st) {
  return gkr_sum_call(1, eq, num, den, lambda, partial, n_terms, n_blocks, st);
}
int vx_gkr_sum_poly_logup_mult(const uint32_t* eq, const uint32_t* num, const uint32_t* den, const uint32_t* lambda,
                               uint32_t* partial, uint32_t n_terms, uint32_t* n_blocks, vx_stream_t st) {
  return gkr_sum_call(2, eq, num, den, lambda, partial, n_terms, n_blocks, st);
}
int vx_gkr_sum_poly_logup_singles(const uint32_t* eq, const uint32_t* den, const uint32_t* lambda, uint32_t* partial,
                                  uint32_t n_terms, uint32_t* n_blocks, vx_stream_t st) {
  return gkr_sum_call(3, eq, nullptr, den, lambda, partial, n_terms, n_blocks, st);
}

// ------------------------------------------------------- batched columns
struct vx_batch {
  BatchCommitter impl;
  vx_batch(const uint32_t* l, uint32_t n, uint32_t b, bool br, hipStream_t s) : impl(l, n, b, br, s) {}
};

int vx_batch_create(const uint32_t* h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, int bit_reverse,
                    vx_stream_t st, vx_batch** out) {
  return guard([&] {
    require(out != nullptr, "batch_create: null out");
    *out = nullptr;
    *out = new vx_batch(h_log_sizes, n_cols, blowup_bits, bit_reverse != 0, S(st));
  });
}

int vx_batch_destroy(vx_batch* b) {
  return guard([&] { delete b; });
}

int vx_batch_column(const vx_batch* b, uint32_t col, int which, uint32_t** d_ptr, uint64_t* words) {
  return guard([&] {
    require(b && d_ptr && words, "null argument");
    *d_ptr = b->impl.column(col, which);
    *words = 1ull << (b->impl.log_size(col) + (which == 2 ? b->impl.blowup() : 0));
  });
}

int vx_batch_enqueue(vx_batch* b, int timed) {
  return guard([&] {
    require(b != nullptr, "null batch");
    b->impl.enqueue(timed != 0);
  });
}

int vx_batch_wait(vx_batch* b, uint32_t h_root[8]) {
  return guard([&] {
    require(b != nullptr, "null batch");
    b->impl.wait(h_root);
  });
}

int vx_batch_stage_times(vx_batch* b, float out[4]) {
  return guard([&] {
    require(b && out, "null argument");
    b->impl.stage_times(out);
  });
}

uint64_t vx_batch_device_bytes(const vx_batch* b) { return b ? b->impl.device_bytes() : 0; }

int vx_m31_fill_random(uint32_t* d, uint32_t log_n, uint64_t seed, uint32_t col, vx_stream_t st) {
  return guard([&] {
    require(d != nullptr && log_n >= 2 && log_n <= 31, "fill_random: bad arguments");
    // column `col` of the synthetic AIR trace generator with this seed
    uint32_t* sw = nullptr;
    uint32_t words[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    DeviceBuf dsw(16);
    VX_HIP(hipMemcpy(dsw.p, words, 8, hipMemcpyHostToDevice));
    sw = dsw.u32();
    d2_trace_col(sw, d, col, log_n, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_circle_ntt_lde_batch(const uint32_t* const* h_in, uint32_t* const* h_coef, uint32_t* const* h_eval,
                            const uint32_t* h_log_sizes, uint32_t n_cols, uint32_t blowup_bits, vx_stream_t st) {
  return guard([&] {
    require(h_in && h_coef && h_eval && h_log_sizes && n_cols >= 1, "lde_batch: null argument");
    hipStream_t s = S(st);
    std::vector<std::shared_ptr<const ProverTables>> keep;
    std::vector<LdeBatchCol> cols(n_cols);
    for (uint32_t c = 0; c < n_cols; c++) {
      require(h_log_sizes[c] >= 14 && h_log_sizes[c] <= 24 && h_log_sizes[c] + blowup_bits <= 31 &&
                  blowup_bits >= 1,
              "lde_batch: log sizes must be 14..24");
      keep.push_back(ProverTables::get(h_log_sizes[c], blowup_bits, s));
      cols[c] = {h_log_sizes[c], h_in[c], h_coef[c], h_eval[c], keep.back()->trace_itw->ref,
                 keep.back()->eval_tw->ref};
    }
    std::vector<uint8_t> jobs;
    LdeBatchStages stg{};
    lde_batch_build(cols.data(), n_cols, blowup_bits, jobs, stg);
    DeviceBuf dj(jobs.size());
    VX_HIP(hipMemcpy(dj.p, jobs.data(), jobs.size(), hipMemcpyHostToDevice));
    lde_batch_launch(static_cast<const uint8_t*>(dj.p), stg, s);
    VX_HIP(hipStreamSynchronize(s));
  });
}

int vx_bit_reverse_m31_batch(const uint32_t* const* h_src, uint32_t* const* h_dst, const uint32_t* h_log_sizes,
                             uint32_t n_cols, vx_stream_t st) {
  return guard([&] {
    require(h_src && h_dst && h_log_sizes && n_cols >= 1, "bit_reverse_batch: null argument");
    std::vector<uint8_t> jobs;
    uint32_t count = 0;
    uint64_t tiles = 0;
    bitrev_batch_build(h_src, h_dst, h_log_sizes, n_cols, jobs, count, tiles);
    DeviceBuf dj(jobs.size());
    VX_HIP(hipMemcpy(dj.p, jobs.data(), jobs.size(), hipMemcpyHostToDevice));
    bitrev_batch_launch(static_cast<const uint8_t*>(dj.p), count, tiles, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

static LeafSrc lifted_src(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                          uint32_t lifting_log, DeviceBuf& table) {
  require(h_cols && h_log_sizes && n_cols >= 1 && lifting_log >= 1 && lifting_log <= 31, "lifted: bad arguments");
  std::vector<LiftCol> lc(n_cols);
  for (uint32_t c = 0; c < n_cols; c++) {
    require(h_cols[c] && h_log_sizes[c] >= 1 && h_log_sizes[c] <= lifting_log, "lifted: column larger than the tree");
    require(c == 0 || h_log_sizes[c] >= h_log_sizes[c - 1], "lifted: columns must be sorted by ascending size");
    lc[c] = {h_cols[c], h_log_sizes[c], 0};
  }
  table = DeviceBuf(sizeof(LiftCol) * n_cols);
  VX_HIP(hipMemcpy(table.p, lc.data(), sizeof(LiftCol) * n_cols, hipMemcpyHostToDevice));
  LeafSrc L{};
  L.kind = LEAF_LIFTED;
  L.lift = static_cast<const LiftCol*>(table.p);
  L.lift_n = n_cols;
  L.lift_log = lifting_log;
  return L;
}

int vx_build_leaves_lifted(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                           uint32_t lifting_log_size, uint32_t* d_hashes, vx_stream_t st) {
  return guard([&] {
    require(d_hashes != nullptr, "lifted: null output");
    DeviceBuf table;
    LeafSrc L = lifted_src(h_cols, h_log_sizes, n_cols, lifting_log_size, table);
    merkle_lifted_leaves(L, d_hashes, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_merkle_root_lifted(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                          uint32_t* d_root, vx_stream_t st) {
  return guard([&] {
    require(d_root != nullptr && h_log_sizes != nullptr && n_cols >= 1, "lifted: null argument");
    const uint32_t lg = h_log_sizes[n_cols - 1];
    require(lg >= 4, "lifted: the largest column needs >= 16 rows");
    DeviceBuf table;
    LeafSrc L = lifted_src(h_cols, h_log_sizes, n_cols, lg, table);
    // only the levels from the tile size up are stored (the tile roots and above)
    const uint32_t tile = merkle_tile_log(L, lg);
    TreeLevels T{};
    uint64_t w = 0;
    T.log_n = lg;
    T.lo = tile;
    for (uint32_t l = T.lo; l <= lg; l++) {
      T.off[l] = w;
      w += 8ull << (lg - l);
    }
    DeviceBuf tree(4 * w);
    T.base = tree.u32();
    RootHook hook{};
    hook.root_out = d_root;
    merkle_commit(L, T, tile, hook, S(st), nullptr, nullptr);
    VX_HIP(hipStreamSynchronize(S(st)));
  });
}

int vx_merkle_root_lifted_noalloc(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                                  uint32_t* d_root, vx_stream_t st) {
  return guard([&] { merkle_root_lifted_noalloc(h_cols, h_log_sizes, n_cols, d_root, S(st)); });
}

// ------------------------------------------------------------ Cairo quotient
// The 34 trace-column pointers arrive as the reference's uploaded device
// array (read back after the stream's earlier work); interaction columns,
// outputs, alphas, Z_H^-1, transition factors and challenges are device
// pointers as in cuda_cairo_quotient.
static CairoArgs cairo_args(const uint32_t* const* d_trace_cols, const uint32_t* const* inter, int n_inter,
                            uint32_t* const out[4], const uint32_t* alpha, const uint32_t* vh, const uint32_t* tf,
                            const uint32_t* chal, hipStream_t s) {
  require(d_trace_cols && alpha && vh && tf && chal, "cairo_quotient: null argument");
  std::vector<uint32_t*> cols = fetch_ptrs(d_trace_cols, CA_COLS, s);
  CairoArgs a{};
  for (uint32_t c = 0; c < CA_COLS; c++) {
    require(cols[c] != nullptr, "cairo_quotient: null trace column");
    a.col[c] = cols[c];
  }
  for (int k = 0; k < n_inter; k++) {
    require(inter[k] != nullptr, "cairo_quotient: null interaction column");
    a.inter[k] = inter[k];
  }
  for (int k = 0; k < 4; k++) {
    require(out[k] != nullptr, "cairo_quotient: null output column");
    a.out[k] = out[k];
  }
  a.alpha = alpha;
  a.vh_inv = vh;
  a.trans = tf;
  a.chal = chal;
  return a;
}

int vx_cairo_quotient(const uint32_t* const* d_trace_cols, const uint32_t* sl0, const uint32_t* sl1,
                      const uint32_t* sl2, const uint32_t* sl3, const uint32_t* t1l0, const uint32_t* t1l1,
                      const uint32_t* t1l2, const uint32_t* t1l3, const uint32_t* t2l0, const uint32_t* t2l1,
                      const uint32_t* t2l2, const uint32_t* t2l3, const uint32_t* t3l0, const uint32_t* t3l1,
                      const uint32_t* t3l2, const uint32_t* t3l3, const uint32_t* sr0, const uint32_t* sr1,
                      const uint32_t* sr2, const uint32_t* sr3, const uint32_t* u1r0, const uint32_t* u1r1,
                      const uint32_t* u1r2, const uint32_t* u1r3, const uint32_t* u2r0, const uint32_t* u2r1,
                      const uint32_t* u2r2, const uint32_t* u2r3, const uint32_t* sd0, const uint32_t* sd1,
                      const uint32_t* sd2, const uint32_t* sd3, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                      uint32_t* out3, const uint32_t* d_alpha, const uint32_t* d_vh_inv, const uint32_t* d_trans,
                      const uint32_t* d_challenges, uint32_t n, uint32_t blowup_step, vx_stream_t st) {
  return guard([&] {
    require(n >= 1 && blowup_step <= n, "cairo_quotient: need n >= 1 and blowup_step <= n");
    const uint32_t* inter[32] = {sl0,  sl1,  sl2,  sl3,  t1l0, t1l1, t1l2, t1l3, t2l0, t2l1, t2l2,
                                 t2l3, t3l0, t3l1, t3l2, t3l3, sr0,  sr1,  sr2,  sr3,  u1r0, u1r1,
                                 u1r2, u1r3, u2r0, u2r1, u2r2, u2r3, sd0,  sd1,  sd2,  sd3};
    uint32_t* out[4] = {out0, out1, out2, out3};
    CairoArgs a = cairo_args(d_trace_cols, inter, 32, out, d_alpha, d_vh_inv, d_trans, d_challenges, S(st));
    a.wrap = n;
    a.blowup_step = blowup_step;
    cairo_quotient(a, n, S(st));
  });
}

int vx_cairo_quotient_slab(const uint32_t* const* d_trace_cols, const uint32_t* sl0, const uint32_t* sl1,
                           const uint32_t* sl2, const uint32_t* sl3, const uint32_t* t1l0, const uint32_t* t1l1,
                           const uint32_t* t1l2, const uint32_t* t1l3, const uint32_t* t2l0, const uint32_t* t2l1,
                           const uint32_t* t2l2, const uint32_t* t2l3, const uint32_t* t3l0, const uint32_t* t3l1,
                           const uint32_t* t3l2, const uint32_t* t3l3, const uint32_t* sr0, const uint32_t* sr1,
                           const uint32_t* sr2, const uint32_t* sr3, const uint32_t* u1r0, const uint32_t* u1r1,
                           const uint32_t* u1r2, const uint32_t* u1r3, const uint32_t* u2r0, const uint32_t* u2r1,
                           const uint32_t* u2r2, const uint32_t* u2r3, const uint32_t* sd0, const uint32_t* sd1,
                           const uint32_t* sd2, const uint32_t* sd3, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                           uint32_t* out3, const uint32_t* d_alpha, const uint32_t* d_vh_inv,
                           const uint32_t* d_trans, const uint32_t* d_challenges, uint32_t chunk_n,
                           uint32_t blowup_step, uint32_t chunk_offset, vx_stream_t st) {
  return guard([&] {
    const uint32_t* inter[32] = {sl0,  sl1,  sl2,  sl3,  t1l0, t1l1, t1l2, t1l3, t2l0, t2l1, t2l2,
                                 t2l3, t3l0, t3l1, t3l2, t3l3, sr0,  sr1,  sr2,  sr3,  u1r0, u1r1,
                                 u1r2, u1r3, u2r0, u2r1, u2r2, u2r3, sd0,  sd1,  sd2,  sd3};
    uint32_t* out[4] = {out0, out1, out2, out3};
    CairoArgs a = cairo_args(d_trace_cols, inter, 32, out, d_alpha, d_vh_inv, d_trans, d_challenges, S(st));
    a.g_base = chunk_offset;
    a.o_base = chunk_offset;
    a.blowup_step = blowup_step;
    cairo_quotient(a, chunk_n, S(st));
  });
}

int vx_cairo_quotient_chunk(const uint32_t* const* d_trace_cols, const uint32_t* sl0, const uint32_t* sl1,
                            const uint32_t* sl2, const uint32_t* sl3, const uint32_t* sr0, const uint32_t* sr1,
                            const uint32_t* sr2, const uint32_t* sr3, const uint32_t* sd0, const uint32_t* sd1,
                            const uint32_t* sd2, const uint32_t* sd3, uint32_t* out0, uint32_t* out1,
                            uint32_t* out2, uint32_t* out3, const uint32_t* d_alpha, const uint32_t* d_vh_inv,
                            const uint32_t* d_trans, const uint32_t* d_challenges, uint32_t offset, uint32_t chunk_n,
                            uint32_t global_n, uint32_t blowup_step, vx_stream_t st) {
  return guard([&] {
    require((uint64_t)offset + chunk_n <= global_n && blowup_step <= global_n,
            "cairo_quotient_chunk: rows beyond global_n");
    const uint32_t* inter[12] = {sl0, sl1, sl2, sl3, sr0, sr1, sr2, sr3, sd0, sd1, sd2, sd3};
    uint32_t* out[4] = {out0, out1, out2, out3};
    CairoArgs a = cairo_args(d_trace_cols, inter, 12, out, d_alpha, d_vh_inv, d_trans, d_challenges, S(st));
    a.col_base = offset;
    a.wrap = global_n;
    a.g_base = offset;
    a.blowup_step = blowup_step;
    a.chunked = 1;
    cairo_quotient(a, chunk_n, S(st));
  });
}

int vx_cairo_quotient_commit(const uint32_t* const* d_trace_cols, const uint32_t* const* h_inter, uint32_t* out0,
                             uint32_t* out1, uint32_t* out2, uint32_t* out3, const uint32_t* d_alpha,
                             const uint32_t* d_vh_inv, const uint32_t* d_trans, const uint32_t* d_challenges,
                             uint32_t log_n, uint32_t blowup_step, uint32_t store_lo, uint32_t* d_levels,
                             uint32_t* d_root, vx_stream_t st) {
  return guard([&] {
    require(h_inter != nullptr && d_root != nullptr, "cairo_quotient_commit: null argument");
    require(log_n >= 2 && log_n <= 31 && blowup_step <= (1ull << log_n), "cairo_quotient_commit: bad sizes");
    const TreeLevels T = caller_levels(d_levels, log_n, store_lo);
    uint32_t* out[4] = {out0, out1, out2, out3};
    CairoArgs a = cairo_args(d_trace_cols, h_inter, 32, out, d_alpha, d_vh_inv, d_trans, d_challenges, S(st));
    a.wrap = 1ull << log_n;
    a.blowup_step = blowup_step;
    DeviceBuf dargs(sizeof(CairoArgs));
    VX_HIP(hipMemcpy(dargs.p, &a, sizeof(CairoArgs), hipMemcpyHostToDevice));
    LeafSrc L{};
    L.kind = LEAF_CAIRO_QUOTIENT;
    L.cairo = static_cast<const CairoArgs*>(dargs.p);
    for (int c = 0; c < 4; c++) L.out[c] = out[c];
    RootHook hook{};
    hook.root_out = d_root;
    merkle_commit(L, T, store_lo, hook, S(st));
    VX_HIP(hipStreamSynchronize(S(st)));  // the argument block outlives the kernels
  });
}

int vx_compute_vanishing_inv(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* d_out, uint32_t log_eval,
                             uint32_t log_n, vx_stream_t st) {
  return guard([&] { cairo_vanishing(ix, iy, sx, sy, d_out, log_eval, log_n, true, S(st)); });
}

int vx_compute_vanishing(uint32_t ix, uint32_t iy, uint32_t sx, uint32_t sy, uint32_t* d_out, uint32_t log_eval,
                         uint32_t log_n, vx_stream_t st) {
  return guard([&] { cairo_vanishing(ix, iy, sx, sy, d_out, log_eval, log_n, false, S(st)); });
}

int vx_axpy_m31(uint32_t scalar, const uint32_t* d_x, uint32_t* d_y, uint32_t n, vx_stream_t st) {
  return guard([&] { axpy_m31(scalar, d_x, d_y, n, S(st)); });
}

// ------------------------------------------------- bytecode constraint eval
static constexpr size_t BC_JIT_MAX_INS = 50000;  // longer programs are interpreted
// The program is read back to the host (it is small), checked and renamed
// to slots, then evaluated; the call returns once the accumulators are
// written (its temporary device buffers are released at return).
static int bytecode_call(const uint32_t* d_bytecode, uint32_t n_words, const uint32_t* const* d_trace_cols,
                         uint32_t n_trace_cols, uint32_t n_rows, uint32_t trace_n_rows,
                         const uint32_t* d_coeff, const uint32_t* d_denom_inv, uint32_t* a0, uint32_t* a1,
                         uint32_t* a2, uint32_t* a3, vx_stream_t st) {
  return guard([&] {
    require(n_rows >= 1 && pow2(n_rows) && trace_n_rows >= 1 && pow2(trace_n_rows) && trace_n_rows <= n_rows,
            "bytecode eval: n_rows and trace_n_rows must be powers of two, trace_n_rows <= n_rows");
    require(n_words == 0 || d_bytecode != nullptr, "bytecode eval: null program");
    require(d_denom_inv && a0 && a1 && a2 && a3, "bytecode eval: null denominators or accumulators");
    require(n_trace_cols == 0 || d_trace_cols != nullptr, "bytecode eval: null column table");
    hipStream_t s = S(st);
    std::vector<uint32_t> words(n_words);
    if (n_words) {
      VX_HIP(hipMemcpyAsync(words.data(), d_bytecode, 4ull * n_words, hipMemcpyDeviceToHost, s));
      VX_HIP(hipStreamSynchronize(s));
    }
    const BcProgram p = bc_compile(words.data(), n_words, n_trace_cols);
    require(!p.uses_offsets || n_rows > trace_n_rows,
            "bytecode eval: row offsets need an evaluation domain larger than the trace domain");
    require(p.n_constraints == 0 || d_coeff != nullptr, "bytecode eval: null constraint coefficients");
    uint32_t* acc[4] = {a0, a1, a2, a3};
    // compiled tier unless switched off or the program is too long to
    // compile in reasonable time; the interpreter takes the rest
    if (bytecode_jit_enabled() && p.ins.size() <= BC_JIT_MAX_INS) {
      bytecode_jit_eval(words.data(), n_words, d_trace_cols, n_trace_cols, n_rows, trace_n_rows, d_coeff, d_denom_inv,
                        acc, s);
      VX_HIP(hipStreamSynchronize(s));
      return;
    }
    DeviceBuf ins(16 * (p.ins.size() ? p.ins.size() : 1)), kon(4 * (p.consts.size() ? p.consts.size() : 4));
    DeviceBuf tr(4 * (p.trace.size() ? p.trace.size() : 3));
    DeviceBuf ws(bytecode_ws_bytes(p, n_rows));
    if (!p.ins.empty()) VX_HIP(hipMemcpyAsync(ins.p, p.ins.data(), 16 * p.ins.size(), hipMemcpyHostToDevice, s));
    if (!p.consts.empty())
      VX_HIP(hipMemcpyAsync(kon.p, p.consts.data(), 4 * p.consts.size(), hipMemcpyHostToDevice, s));
    // the device moves natural indices by offset * 2^(eval_log - trace_log - 1)
    // modulo the half domain (stwo's offset_bit_reversed_circle_domain_index)
    std::vector<uint32_t> tr_h(p.trace);
    const uint32_t e = log2u(n_rows), d = log2u(trace_n_rows);
    for (size_t i = 1; i < tr_h.size(); i += 3) {
      const int32_t off = (int32_t)tr_h[i];
      if (off == 0) continue;
      const int64_t half = (int64_t)1 << (e - 1);
      int64_t st = ((int64_t)off * ((int64_t)1 << (e - d - 1))) % half;
      if (st < 0) st += half;
      tr_h[i] = (uint32_t)st;
    }
    if (!tr_h.empty()) VX_HIP(hipMemcpyAsync(tr.p, tr_h.data(), 4 * tr_h.size(), hipMemcpyHostToDevice, s));
    bytecode_constraint_eval(p, ins.u32(), kon.u32(), tr.u32(), d_trace_cols, n_rows, trace_n_rows, d_coeff,
                             d_denom_inv, acc, ws.p ? ws.u32() : nullptr, s);
    VX_HIP(hipStreamSynchronize(s));
  });
}

int vx_bytecode_constraint_eval(const uint32_t* d_bytecode, uint32_t n_words, const uint32_t* const* d_trace_cols,
                                const uint32_t* d_trace_col_sizes, uint32_t n_trace_cols, uint32_t n_rows,
                                uint32_t trace_n_rows, const uint32_t* d_random_coeff_powers,
                                const uint32_t* d_denom_inv, uint32_t log_expand, uint32_t* accum0,
                                uint32_t* accum1, uint32_t* accum2, uint32_t* accum3, uint32_t n_registers,
                                vx_stream_t stream) {
  (void)d_trace_col_sizes;  // reserved in the reference as well
  (void)log_expand;         // implied by n_rows / trace_n_rows
  (void)n_registers;        // the slot count comes from the program itself
  return bytecode_call(d_bytecode, n_words, d_trace_cols, n_trace_cols, n_rows, trace_n_rows, d_random_coeff_powers,
                       d_denom_inv, accum0, accum1, accum2, accum3, stream);
}

int vx_warp_bytecode_constraint_eval(const uint32_t* d_bytecode, uint32_t n_words,
                                     const uint32_t* const* d_trace_cols, const uint32_t* d_trace_col_sizes,
                                     uint32_t n_trace_cols, uint32_t n_rows, uint32_t trace_n_rows,
                                     const uint32_t* d_random_coeff_powers, const uint32_t* d_denom_inv,
                                     uint32_t log_expand, uint32_t* accum0, uint32_t* accum1, uint32_t* accum2,
                                     uint32_t* accum3, uint32_t n_registers, vx_stream_t stream) {
  return vx_bytecode_constraint_eval(d_bytecode, n_words, d_trace_cols, d_trace_col_sizes, n_trace_cols, n_rows,
                                     trace_n_rows, d_random_coeff_powers, d_denom_inv, log_expand, accum0, accum1,
                                     accum2, accum3, n_registers, stream);
}

int vx_bytecode_jit_source(const uint32_t* h_bytecode, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                           uint32_t trace_log, char* out, size_t cap, size_t* out_len) {
  return guard([&] {
    require(eval_log >= 1 && eval_log <= 31 && trace_log <= eval_log, "bytecode JIT: bad domain sizes");
    (void)bc_compile(h_bytecode, n_words, n_trace_cols);  // the same checks as an evaluation
    size_t n_ops = 0;
    const std::string src = bc_jit_source(h_bytecode, n_words, n_trace_cols, eval_log, trace_log, &n_ops);
    if (out_len) *out_len = src.size();
    if (out) {
      require(cap > src.size(), "bytecode JIT: output buffer too small");
      std::memcpy(out, src.data(), src.size());
      out[src.size()] = 0;
    }
  });
}

int vx_bytecode_jit_compile(const uint32_t* h_bytecode, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                            uint32_t trace_log, size_t* code_bytes) {
  return guard([&] {
    require(eval_log >= 1 && eval_log <= 31 && trace_log <= eval_log, "bytecode JIT: bad domain sizes");
    (void)bc_compile(h_bytecode, n_words, n_trace_cols);
    const size_t n = bytecode_jit_code_bytes(h_bytecode, n_words, n_trace_cols, eval_log, trace_log);
    if (code_bytes) *code_bytes = n;
  });
}

int vx_bytecode_set_jit(int on) {
  return guard([&] { bytecode_jit_enable(on != 0); });
}

int vx_bytecode_slots(const uint32_t* h_bytecode, uint32_t n_words, uint32_t n_trace_cols, uint32_t* m31_slots,
                      uint32_t* qm31_slots, uint32_t* n_constraints) {
  return guard([&] {
    const BcProgram p = bc_compile(h_bytecode, n_words, n_trace_cols);
    if (m31_slots) *m31_slots = p.m31_slots;
    if (qm31_slots) *qm31_slots = p.qm31_slots;
    if (n_constraints) *n_constraints = p.n_constraints;
  });
}

// ------------------------------------------------------------------ serving
struct vx_server {
  ProofServer impl;
  vx_server(uint32_t n, uint64_t budget, bool graphs) : impl(n, budget, graphs) {}
};

static ProofRequest to_request(const vx_request* r) {
  require(r != nullptr, "server: null request");
  ProofRequest q{};
  q.air = r->air;
  q.log_n = r->log_n;
  q.blowup_bits = r->blowup_bits;
  q.n_queries = r->n_queries;
  q.pow_bits = r->pow_bits;
  q.a = r->a;
  q.b = r->b;
  q.n_cols = r->n_cols;
  q.n_cons = r->n_cons;
  q.seed = r->seed;
  q.air_seed = r->air_seed;
  return q;
}

int vx_server_create(uint32_t n_slots, uint64_t mem_budget_bytes, int use_graphs, vx_server** out) {
  return guard([&] {
    require(out != nullptr, "server_create: null out");
    *out = nullptr;
    *out = new vx_server(n_slots, mem_budget_bytes, use_graphs != 0);
  });
}

int vx_server_destroy(vx_server* s) {
  return guard([&] { delete s; });
}

int vx_server_prepare(vx_server* s, const vx_request* r) {
  return guard([&] {
    require(s != nullptr, "null server");
    s->impl.prepare(to_request(r));
  });
}

int vx_server_submit(vx_server* s, const vx_request* r, uint8_t* out, size_t out_cap, uint64_t* ticket) {
  return guard([&] {
    require(s && ticket, "null argument");
    *ticket = s->impl.submit(to_request(r), out, out_cap);
  });
}

int vx_server_wait(vx_server* s, uint64_t ticket, size_t* out_len, double times[3]) {
  return guard([&] {
    require(s != nullptr, "null server");
    JobTimes t{};
    std::string err;
    bool ok = s->impl.wait(ticket, out_len, &t, err);
    if (times) {
      times[0] = t.submit_s;
      times[1] = t.start_s;
      times[2] = t.done_s;
    }
    if (!ok) throw Error(ERR_ARG, err);
  });
}

int vx_server_done(vx_server* s, uint64_t ticket) {
  int done = 0;
  int rc = guard([&] {
    require(s != nullptr, "null server");
    done = s->impl.done(ticket) ? 1 : 0;
  });
  return rc ? rc : done;
}

uint64_t vx_server_resident_bytes(const vx_server* s) { return s ? s->impl.resident_bytes() : 0; }

int vx_server_stats(const vx_server* s, uint64_t out[3]) {
  return guard([&] {
    require(s && out, "null argument");
    const ServerStats st = s->impl.stats();
    out[0] = st.created;
    out[1] = st.evicted;
    out[2] = st.deferred;
  });
}

int vx_server_run_trace(vx_server* s, const vx_request* reqs, const double* arrival_s, uint32_t n,
                        uint8_t* const* outs, const size_t* caps, size_t* lens, double* times) {
  return guard([&] {
    require(s && reqs && arrival_s && outs && caps && lens && times, "null argument");
    for (uint32_t i = 1; i < n; i++) require(arrival_s[i] >= arrival_s[i - 1], "run_trace: arrivals must be sorted");
    std::vector<uint64_t> tickets(n);
    const double t0 = s->impl.now();
    for (uint32_t i = 0; i < n; i++) {
      const double due = t0 + arrival_s[i];
      for (;;) {  // sleep through long gaps, spin the last 200 us
        double left = due - s->impl.now();
        if (left <= 0) break;
        if (left > 3e-4) std::this_thread::sleep_for(std::chrono::duration<double>(left - 2e-4));
      }
      tickets[i] = s->impl.submit(to_request(&reqs[i]), outs[i], caps[i]);
    }
    std::string first_err;
    for (uint32_t i = 0; i < n; i++) {
      JobTimes t{};
      std::string err;
      bool ok = s->impl.wait(tickets[i], &lens[i], &t, err);
      if (!ok && first_err.empty()) first_err = "request " + std::to_string(i) + ": " + err;
      times[3 * i + 0] = t.submit_s - t0;
      times[3 * i + 1] = t.start_s - t0;
      times[3 * i + 2] = t.done_s - t0;
    }
    if (!first_err.empty()) throw Error(ERR_ARG, first_err);
  });
}

}  // extern "C"
