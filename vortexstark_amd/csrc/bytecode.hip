// Bytecode constraint evaluation on gfx950 (cuda/constraint_eval.cu:97-451,
// cuda/constraint_eval_warp.cu): every row of the evaluation domain runs
// the same straight-line program; the accumulator columns gain
// denom_inv[row >> trace_log] * sum_k coeff_k * constraint_k(row).
//
// MI355X design. The program arrives renamed to typed slots (bytecode.cpp),
// so one row's register file is the program's peak number of live values.
// Instructions are 16-byte records read with scalar loads: the opcode,
// operand slots and constants are wave-uniform, the dispatch is a scalar
// branch, and each lane only computes its row. When the register file stays
// small, a prologue loads every distinct (column, offset) the program reads
// into slots, eight loads in flight at a time; otherwise trace values are
// read from the columns at each use (L1/L2 hits after the first), which
// keeps the register file, and so the wavefronts per CU, what the computed
// values need. Constants are operands.
// A wavefront's register file lives in LDS when it is at most 128 words per
// row (32 KiB per 64-row workgroup, 5 waves per CU); larger programs keep it
// in a global workspace laid out slot-major ([slot][row]), so every operand
// access of a wavefront is one contiguous 256-byte (QM31: 1 KiB) run.
#include "vx_bytecode.h"
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t BC_LDS_ROWS = 64;       // one wavefront per workgroup
static constexpr uint32_t BC_LDS_MAX_WORDS = 128;  // words per row kept in LDS
static constexpr uint32_t BC_WS_ROWS = 256;       // workgroup of the workspace kernel
static constexpr uint64_t BC_WS_BYTES = 1ull << 30;  // workspace per launch

struct BcArgs {
  const BcIns* ins;
  uint32_t n_ins;
  const uint32_t* consts;  // QM31 constants, 4 words each
  const uint32_t* trace;   // (column, step mod half domain, slot) of each distinct trace load
  uint32_t n_trace;
  uint32_t remat;          // trace values have no slots: read at each use
  const uint32_t* const* cols;
  const uint32_t* coeff;   // QM31 per constraint
  const uint32_t* denom_inv;
  uint32_t* acc[4];
  uint32_t n_rows, eval_log, trace_log;
  uint32_t m31_slots;      // LDS: QM31 slots start after the M31 ones
  uint32_t row0;           // first row of this launch
  uint32_t* ws_m;          // workspace: M31 slots [slot][chunk], then QM31 slots
  uint4* ws_q;
  uint32_t chunk;          // rows per workspace slot
};

template <bool LDS>
__global__ __launch_bounds__(LDS ? BC_LDS_ROWS : BC_WS_ROWS) void k_bc_eval(BcArgs a) {
  extern __shared__ uint32_t bc_lds[];
  const uint32_t t = threadIdx.x;
  const uint32_t lr = blockIdx.x * blockDim.x + t;  // row within this launch
  const uint32_t r = a.row0 + lr;
  const bool live = r < a.n_rows;
  const uint32_t row = live ? r : a.n_rows - 1;  // idle lanes read a valid row and store nothing

  auto ld_m = [&](uint32_t s) -> uint32_t {
    if constexpr (LDS) return bc_lds[s * BC_LDS_ROWS + t];
    else return as_global(a.ws_m)[(uint64_t)s * a.chunk + lr];
  };
  auto st_m = [&](uint32_t s, uint32_t v) {
    if constexpr (LDS) bc_lds[s * BC_LDS_ROWS + t] = v;
    else ((__attribute__((address_space(1))) uint32_t*)a.ws_m)[(uint64_t)s * a.chunk + lr] = v;
  };
  auto ld_q = [&](uint32_t s) -> qm31 {
    uint4 v;
    if constexpr (LDS) v = reinterpret_cast<const uint4*>(bc_lds + a.m31_slots * BC_LDS_ROWS)[s * BC_LDS_ROWS + t];
    else v = ld4_global(reinterpret_cast<const uint32_t*>(a.ws_q + (uint64_t)s * a.chunk + lr));
    return {{v.x, v.y, v.z, v.w}};
  };
  auto st_q = [&](uint32_t s, const qm31& q) {
    const uint4 v = make_uint4(q.v[0], q.v[1], q.v[2], q.v[3]);
    if constexpr (LDS) {
      reinterpret_cast<uint4*>(bc_lds + a.m31_slots * BC_LDS_ROWS)[s * BC_LDS_ROWS + t] = v;
    } else {
      __attribute__((address_space(1))) uint32_t* g =
          (__attribute__((address_space(1))) uint32_t*)(a.ws_q + (uint64_t)s * a.chunk + lr);
      g[0] = v.x;
      g[1] = v.y;
      g[2] = v.z;
      g[3] = v.w;
    }
  };
  auto konst = [&](uint32_t k) -> qm31 {
    const uint4 c = ld_uniform(reinterpret_cast<const uint4*>(a.consts) + k);
    return {{c.x, c.y, c.z, c.w}};
  };

  // A row offset moves the natural index by a step s (in the first half of
  // the domain, back by s in the conjugate half) modulo the half size; the
  // host stores s mod half, so a move is an add and a mask.
  const uint32_t hmask = a.eval_log ? (1u << (a.eval_log - 1)) - 1u : 0u;
  const uint32_t nat = a.eval_log ? __builtin_bitreverse32(row) >> (32u - a.eval_log) : 0u;
  const uint32_t second = nat & ~hmask;  // the conjugate half
  auto trace_val = [&](uint32_t t) -> uint32_t {
    const uint32_t* e = a.trace + 3 * t;
    const uint32_t cix = ld_uniform(e), step = ld_uniform(e + 1);
    const uint32_t* col = ld_uniform(a.cols + cix);
    if (col == nullptr) return 0u;  // a null column reads as zero
    uint32_t rr = row;
    if (step) {
      const uint32_t moved = second ? ((nat - step) & hmask) | second : (nat + step) & hmask;
      rr = __builtin_bitreverse32(moved) >> (32u - a.eval_log);
    }
    return as_global(col)[rr];
  };
  // Prologue: trace values that have slots, eight loads in flight at a time.
  if (!a.remat) {
    for (uint32_t t0 = 0; t0 < a.n_trace; t0 += 8) {
      uint32_t v[8];
#pragma unroll
      for (uint32_t j = 0; j < 8; j++) v[j] = t0 + j < a.n_trace ? trace_val(t0 + j) : 0u;
#pragma unroll
      for (uint32_t j = 0; j < 8; j++)
        if (t0 + j < a.n_trace) st_m(ld_uniform(a.trace + 3 * (t0 + j) + 2), v[j]);
    }
  }

  // operand of kind k as M31 (component 0) or as QM31 ((v, 0, 0, 0) for an
  // M31 source); the kind is wave-uniform
  auto get1 = [&](uint32_t k, uint32_t ix) -> uint32_t {
    if (k == BC_K_M) return ld_m(ix);
    if (k == BC_K_Q) return ld_q(ix).v[0];
    if (k == BC_K_TRACE) return trace_val(ix);
    return konst(ix).v[0];
  };
  auto getq = [&](uint32_t k, uint32_t ix) -> qm31 {
    if (k == BC_K_Q) return ld_q(ix);
    if (k == BC_K_CONST) return konst(ix);
    return {{k == BC_K_M ? ld_m(ix) : trace_val(ix), 0u, 0u, 0u}};
  };

  qm31 acc = {{0u, 0u, 0u, 0u}};
  BcIns nxt = a.n_ins ? ld_uniform(a.ins) : BcIns{0u, 0u, 0u, 0u};
  for (uint32_t pc = 0; pc < a.n_ins; pc++) {
    const BcIns I = nxt;  // the next record is in flight while this one runs
    if (pc + 1 < a.n_ins) nxt = ld_uniform(a.ins + pc + 1);
    const uint32_t op = I.x >> 24, dst = I.x & 0xFFFFu, sa = I.y >> 16, sb = I.y & 0xFFFFu;
    const uint32_t qa = (I.x >> BC_KA) & 3u, qb = (I.x >> BC_KB) & 3u;
    switch (op) {  // trace loads and constants have no instruction of their own
      case BC_ADD:
        st_m(dst, m31_add(get1(qa, sa), get1(qb, sb)));
        break;
      case BC_SUB:
        st_m(dst, m31_sub(get1(qa, sa), get1(qb, sb)));
        break;
      case BC_MUL:
        st_m(dst, m31_mul(get1(qa, sa), get1(qb, sb)));
        break;
      case BC_NEG:
        st_m(dst, m31_neg(get1(qa, sa)));
        break;
      case BC_ADD_CONST:
        st_m(dst, m31_add(get1(qa, sa), I.z));
        break;
      case BC_MUL_CONST:
        st_m(dst, m31_mul(get1(qa, sa), I.z));
        break;
      case BC_WIDE_ADD:
        st_q(dst, qm_add(getq(qa, sa), getq(qb, sb)));
        break;
      case BC_WIDE_SUB:
        st_q(dst, qm_sub(getq(qa, sa), getq(qb, sb)));
        break;
      case BC_WIDE_MUL:
        st_q(dst, qm_mul(getq(qa, sa), getq(qb, sb)));
        break;
      case BC_WIDE_NEG:
        st_q(dst, qm_sub(qm31{{0u, 0u, 0u, 0u}}, getq(qa, sa)));
        break;
      case BC_WIDE_ADD_CONST:
        st_q(dst, qm_add(getq(qa, sa), konst(I.z)));
        break;
      case BC_WIDE_MUL_CONST:
        st_q(dst, qm_mul(getq(qa, sa), konst(I.z)));
        break;
      case BC_WIDE_ADD_BASE: {
        qm31 x = getq(qa, sa);
        x.v[0] = m31_add(x.v[0], get1(qb, sb));
        st_q(dst, x);
        break;
      }
      case BC_WIDE_MUL_BASE:
        st_q(dst, qm_scale(getq(qa, sa), get1(qb, sb)));
        break;
      case BC_BASE_ADD_SECURE_CONST:
        st_q(dst, qm_add(qm31{{get1(qa, sa), 0u, 0u, 0u}}, konst(I.z)));
        break;
      case BC_BASE_MUL_SECURE_CONST:
        st_q(dst, qm_scale(konst(I.z), get1(qa, sa)));
        break;
      case BC_WIDEN:
        st_q(dst, qm31{{get1(qa, sa), 0u, 0u, 0u}});
        break;
      case BC_COMBINE_EF:
        st_q(dst, qm31{{get1(qa, sa), get1(qb, sb), get1((I.x >> BC_KC) & 3u, I.w >> 16),
                        get1((I.x >> BC_KD) & 3u, I.w & 0xFFFFu)}});
        break;
      case BC_ADD_CONSTRAINT: {
        const uint4 c = ld_uniform(reinterpret_cast<const uint4*>(a.coeff) + I.z);
        acc = qm_add(acc, qm_mul(getq(qa, sa), qm31{{c.x, c.y, c.z, c.w}}));
        break;
      }
      default:
        break;  // the host pass admits only the opcodes above
    }
  }
  if (live) {
    const qm31 res = qm_scale(acc, a.denom_inv[row >> a.trace_log]);
#pragma unroll
    for (int c = 0; c < 4; c++) a.acc[c][row] = m31_add(a.acc[c][row], res.v[c]);
  }
}

void bytecode_constraint_eval(const BcProgram& p, const uint32_t* d_ins, const uint32_t* d_consts,
                              const uint32_t* d_trace, const uint32_t* const* d_cols, uint32_t n_rows, uint32_t trace_n_rows,
                              const uint32_t* coeff, const uint32_t* denom_inv, uint32_t* const acc[4],
                              uint32_t* d_ws, hipStream_t s) {
  BcArgs a{};
  a.ins = reinterpret_cast<const BcIns*>(d_ins);
  a.n_ins = (uint32_t)p.ins.size();
  a.consts = d_consts;
  a.trace = d_trace;
  a.n_trace = (uint32_t)(p.trace.size() / 3);
  a.remat = p.remat ? 1u : 0u;
  a.cols = d_cols;
  a.coeff = coeff;
  a.denom_inv = denom_inv;
  for (int c = 0; c < 4; c++) a.acc[c] = acc[c];
  a.n_rows = n_rows;
  a.eval_log = 31u - (uint32_t)__builtin_clz(n_rows);
  a.trace_log = 31u - (uint32_t)__builtin_clz(trace_n_rows);
  a.m31_slots = p.m31_slots;
  const uint32_t words = bytecode_row_words(p);
  if (words <= BC_LDS_MAX_WORDS) {
    const uint32_t grid = (n_rows + BC_LDS_ROWS - 1) / BC_LDS_ROWS;
    hipLaunchKernelGGL(k_bc_eval<true>, dim3(grid), dim3(BC_LDS_ROWS), (size_t)4 * BC_LDS_ROWS * words, s, a);
    VX_HIP(hipGetLastError());
    return;
  }
  const uint32_t chunk = bytecode_ws_rows(p, n_rows);
  a.chunk = chunk;
  a.ws_m = d_ws;
  a.ws_q = reinterpret_cast<uint4*>(d_ws + (uint64_t)p.m31_slots * chunk);
  for (uint64_t r0 = 0; r0 < n_rows; r0 += chunk) {
    a.row0 = (uint32_t)r0;
    const uint64_t rows = n_rows - r0 < chunk ? n_rows - r0 : chunk;
    hipLaunchKernelGGL(k_bc_eval<false>, dim3((uint32_t)((rows + BC_WS_ROWS - 1) / BC_WS_ROWS)), dim3(BC_WS_ROWS), 0,
                       s, a);
    VX_HIP(hipGetLastError());
  }
}

uint32_t bytecode_row_words(const BcProgram& p) { return p.m31_slots + 4 * p.qm31_slots; }

uint32_t bytecode_ws_rows(const BcProgram& p, uint32_t n_rows) {
  const uint64_t per_row = 4ull * bytecode_row_words(p);
  uint64_t chunk = (BC_WS_BYTES / per_row) & ~(uint64_t)(BC_WS_ROWS - 1);
  const uint64_t all = ((uint64_t)n_rows + BC_WS_ROWS - 1) & ~(uint64_t)(BC_WS_ROWS - 1);
  if (chunk > all) chunk = all;
  return (uint32_t)(chunk ? chunk : BC_WS_ROWS);
}

uint64_t bytecode_ws_bytes(const BcProgram& p, uint32_t n_rows) {
  if (bytecode_row_words(p) <= BC_LDS_MAX_WORDS) return 0;
  return 4ull * bytecode_row_words(p) * bytecode_ws_rows(p, n_rows);
}

}  // namespace vx
