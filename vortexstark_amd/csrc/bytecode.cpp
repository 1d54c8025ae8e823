// Host pass of the bytecode constraint evaluator: decode the reference's
// register bytecode (crates/cuda-backend/src/constraint_eval/bytecode.rs:
// 173-323 writes it, cuda/constraint_eval.cu:122-407 reads it), check it,
// and rename its registers to typed slots for the device.
//
// The reference kernel gives every row thread a 1,024-entry QM31 register
// file in local memory. Here the program is straight-line code, so the live
// range of every value is known before the launch: a value lives from the
// instruction that writes it to its last read. Trace loads have no inputs,
// so they are hoisted into a prologue (one value per distinct column and
// offset, all loads of a row in flight together) and constants become
// operands of the instructions that read them. A linear scan assigns each
// value a slot of its type (M31: one word per row, QM31: four), reusing
// slots whose value has died, so a wavefront's register file is as small as
// the program's peak number of live values and fits in LDS.
#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <string>

#include "vx_bytecode.h"
#include "vx_error.h"

namespace vx {

namespace {

enum : uint8_t { T_M31 = 0, T_QM31 = 1 };

struct Dec {
  uint32_t op;
  uint32_t dst;          // virtual register (or the source of ADD_CONSTRAINT)
  uint32_t src[4];       // virtual registers read
  uint32_t n_src;
  uint32_t imm;          // M31 constant or flat column
  int32_t offset;        // LOAD_TRACE
  uint32_t k[4];         // QM31 constant
  bool has_k;
};

uint8_t result_type(uint32_t op) {
  switch (op) {
    case BC_LOAD_CONST:
    case BC_LOAD_TRACE:
    case BC_ADD:
    case BC_SUB:
    case BC_MUL:
    case BC_NEG:
    case BC_ADD_CONST:
    case BC_MUL_CONST:
      return T_M31;
    default:
      return T_QM31;
  }
}

}  // namespace

namespace {

// 1. decode and check
std::vector<Dec> decode(const uint32_t* w, uint32_t n_words, uint32_t n_trace_cols) {
  require(w != nullptr || n_words == 0, "bytecode: null program");
  std::vector<Dec> ops;
  uint32_t pc = 0;
  auto take = [&]() -> uint32_t {
    if (pc >= n_words) throw Error(ERR_ARG, "bytecode: program ends inside an instruction");
    return w[pc++];
  };
  while (pc < n_words) {
    const uint32_t hdr = take();
    Dec d{};
    d.op = hdr >> 24;
    d.dst = hdr & 0xFFFFu;
    switch (d.op) {
      case BC_LOAD_CONST:
        d.imm = take();
        break;
      case BC_LOAD_SECURE_CONST:
        for (auto& x : d.k) x = take();
        d.has_k = true;
        break;
      case BC_LOAD_TRACE: {
        d.imm = take();
        const uint32_t ext = take();
        const uint32_t mag = ext & 0x7FFFFFFFu;
        d.offset = (ext >> 31) ? -(int32_t)mag : (int32_t)mag;
        if (d.imm >= n_trace_cols)
          throw Error(ERR_ARG, "bytecode: trace column " + std::to_string(d.imm) + " out of range");
        break;
      }
      case BC_ADD:
      case BC_SUB:
      case BC_MUL:
      case BC_WIDE_ADD:
      case BC_WIDE_SUB:
      case BC_WIDE_MUL:
      case BC_WIDE_ADD_BASE:
      case BC_WIDE_MUL_BASE: {
        const uint32_t s = take();
        d.src[0] = s >> 16;
        d.src[1] = s & 0xFFFFu;
        d.n_src = 2;
        break;
      }
      case BC_NEG:
      case BC_WIDE_NEG:
      case BC_WIDEN:
        d.src[0] = take() >> 16;
        d.n_src = 1;
        break;
      case BC_ADD_CONST:
      case BC_MUL_CONST:
        d.src[0] = take() >> 16;
        d.n_src = 1;
        d.imm = take();
        break;
      case BC_WIDE_ADD_CONST:
      case BC_WIDE_MUL_CONST:
      case BC_BASE_ADD_SECURE_CONST:
      case BC_BASE_MUL_SECURE_CONST:
        d.src[0] = take() >> 16;
        d.n_src = 1;
        for (auto& x : d.k) x = take();
        d.has_k = true;
        break;
      case BC_COMBINE_EF: {
        const uint32_t s01 = take(), s23 = take();
        d.src[0] = s01 >> 16;
        d.src[1] = s01 & 0xFFFFu;
        d.src[2] = s23 >> 16;
        d.src[3] = s23 & 0xFFFFu;
        d.n_src = 4;
        break;
      }
      case BC_ADD_CONSTRAINT:
        d.src[0] = d.dst;
        d.n_src = 1;
        break;
      default:
        throw Error(ERR_ARG, "bytecode: unknown opcode " + std::to_string(d.op) + " at word " +
                                 std::to_string(pc - 1));
    }
    // constants must be canonical M31 words, as every other input
    const bool imm_const = d.op == BC_LOAD_CONST || d.op == BC_ADD_CONST || d.op == BC_MUL_CONST;
    if ((imm_const && d.imm >= P) || (d.has_k && (d.k[0] >= P || d.k[1] >= P || d.k[2] >= P || d.k[3] >= P)))
      throw Error(ERR_ARG, "bytecode: constant is not a canonical M31 value (word " + std::to_string(pc - 1) + ")");
    ops.push_back(d);
  }

  return ops;
}

// 2-3. values and slots. remat: trace loads take no slot; an operand of
// kind TRACE reads the column where it is used.
BcProgram lower(const std::vector<Dec>& ops, bool remat) {
  // 2. values: every write of a register starts a new value; find each
  // value's type and last read. A trace load names the value of its
  // (column, offset) pair, loaded in the prologue; a constant load makes the
  // register an alias of a constant-pool entry (operand kind CONST).
  const size_t n = ops.size();
  BcProgram p;
  p.remat = remat;
  std::vector<int32_t> cur(1u << 16, -1);  // register -> live value
  std::vector<uint8_t> vtype, vkind;       // kind: BC_K_M, BC_K_Q or BC_K_CONST
  std::vector<uint32_t> vidx;              // constant-pool index of a constant value
  std::vector<size_t> last;
  std::vector<std::array<int32_t, 4>> src_val(n);
  std::vector<int32_t> dst_val(n, -1);
  std::map<std::pair<uint32_t, int32_t>, int32_t> trace_val;  // (column, offset) -> value
  std::vector<int32_t> trace_vals;                             // in table order
  std::map<std::array<uint32_t, 4>, uint32_t> const_ix;
  auto konst = [&](const std::array<uint32_t, 4>& k) {
    auto it = const_ix.find(k);
    if (it != const_ix.end()) return it->second;
    const uint32_t ix = (uint32_t)(p.consts.size() / 4);
    for (uint32_t x : k) p.consts.push_back(x);
    const_ix.emplace(k, ix);
    return ix;
  };
  auto new_value = [&](uint8_t type, uint8_t kind, uint32_t ix, size_t at) {
    vtype.push_back(type);
    vkind.push_back(kind);
    vidx.push_back(ix);
    last.push_back(at);
    return (int32_t)vtype.size() - 1;
  };
  for (size_t i = 0; i < n; i++) {
    const Dec& d = ops[i];
    for (uint32_t s = 0; s < d.n_src; s++) {
      const int32_t v = cur[d.src[s]];
      if (v < 0)
        throw Error(ERR_ARG, "bytecode: register " + std::to_string(d.src[s]) + " read before it is written");
      src_val[i][s] = v;
      last[v] = i;
    }
    if (d.op == BC_ADD_CONSTRAINT) continue;
    int32_t v;
    if (d.op == BC_LOAD_TRACE) {
      auto key = std::make_pair(d.imm, d.offset);
      auto it = trace_val.find(key);
      if (it == trace_val.end()) {
        const uint32_t tix = (uint32_t)(p.trace.size() / 3);
        it = trace_val.emplace(key, new_value(0, remat ? BC_K_TRACE : BC_K_M, tix, 0)).first;
        trace_vals.push_back(it->second);
        p.trace.push_back(d.imm);
        p.trace.push_back((uint32_t)d.offset);
        p.trace.push_back(BC_NO_SLOT);  // slot, below (none when re-read at each use)
        p.uses_offsets = p.uses_offsets || d.offset != 0;
      }
      v = it->second;
    } else if (d.op == BC_LOAD_CONST) {
      v = new_value(0, BC_K_CONST, konst({d.imm, 0u, 0u, 0u}), i);
    } else if (d.op == BC_LOAD_SECURE_CONST) {
      v = new_value(1, BC_K_CONST, konst({d.k[0], d.k[1], d.k[2], d.k[3]}), i);
    } else {
      const uint8_t t = result_type(d.op);
      v = new_value(t, t ? BC_K_Q : BC_K_M, 0, i);
    }
    cur[d.dst] = v;
    dst_val[i] = v;
  }
  if (p.trace.size() / 3 > 0xFFFFu || p.consts.size() / 4 > 0xFFFFu)
    throw Error(ERR_ARG, "bytecode: more than 65,535 distinct trace loads or constants");

  // 3. linear scan: the prologue's values take the first M31 slots; then,
  // per instruction, free the slots of values read for the last time there
  // and give the result a free slot of its type (the device reads every
  // operand before it writes the result, so the two may share a slot)
  std::vector<uint32_t> slot(vtype.size(), 0);
  std::set<uint32_t> free_slots[2];
  uint32_t n_slots[2] = {0, 0};
  auto slotted = [&](int32_t v) { return vkind[v] == BC_K_M || vkind[v] == BC_K_Q; };
  if (!remat) {
    for (size_t t = 0; t < trace_vals.size(); t++) {
      slot[trace_vals[t]] = n_slots[0]++;
      p.trace[3 * t + 2] = slot[trace_vals[t]];
    }
  }
  std::vector<uint8_t> dead(vtype.size(), 0);
  for (size_t i = 0; i < n; i++) {
    const Dec& d = ops[i];
    if (d.op == BC_LOAD_TRACE || d.op == BC_LOAD_CONST || d.op == BC_LOAD_SECURE_CONST) continue;
    BcIns in{};
    uint32_t s_ix[4] = {0, 0, 0, 0}, s_kind[4] = {0, 0, 0, 0};
    for (uint32_t s = 0; s < d.n_src; s++) {
      const int32_t v = src_val[i][s];
      s_ix[s] = slotted(v) ? slot[v] : vidx[v];
      s_kind[s] = vkind[v];
    }
    for (uint32_t s = 0; s < d.n_src; s++) {
      const int32_t v = src_val[i][s];
      if (slotted(v) && last[v] == i && !dead[v]) {  // the same value read twice is freed once
        dead[v] = 1;
        free_slots[vtype[v]].insert(slot[v]);
      }
    }
    uint32_t dslot = 0;
    if (dst_val[i] >= 0) {
      const uint8_t t = vtype[dst_val[i]];
      if (!free_slots[t].empty()) {  // lowest free slot first keeps the live set compact
        dslot = *free_slots[t].begin();
        free_slots[t].erase(free_slots[t].begin());
      } else {
        dslot = n_slots[t]++;
      }
      slot[dst_val[i]] = dslot;
      if (last[dst_val[i]] == i) free_slots[t].insert(dslot);  // never read
    }
    if (n_slots[0] > 0xFFFFu || n_slots[1] > 0xFFFFu) throw Error(ERR_ARG, "bytecode: too many live values");
    in.x = (d.op << 24) | (s_kind[0] << BC_KA) | (s_kind[1] << BC_KB) | (s_kind[2] << BC_KC) |
           (s_kind[3] << BC_KD) | dslot;
    in.y = (s_ix[0] << 16) | s_ix[1];
    switch (d.op) {
      case BC_ADD_CONST:
      case BC_MUL_CONST:
        in.z = d.imm;
        break;
      case BC_ADD_CONSTRAINT:
        in.z = p.n_constraints++;
        break;
      default:
        if (d.has_k) in.z = konst({d.k[0], d.k[1], d.k[2], d.k[3]});
    }
    if (d.op == BC_COMBINE_EF) in.w = (s_ix[2] << 16) | s_ix[3];
    p.ins.push_back(in);
  }
  p.m31_slots = n_slots[0];
  p.qm31_slots = n_slots[1];
  return p;
}

}  // namespace

BcProgram bc_compile(const uint32_t* w, uint32_t n_words, uint32_t n_trace_cols) {
  const std::vector<Dec> ops = decode(w, n_words, n_trace_cols);
  // Trace values in slots are loaded once per row, but they are live from
  // the prologue to their last use; when that makes the register file large
  // (fewer wavefronts per CU to hide latency) they are re-read at each use.
  BcProgram p = lower(ops, false);
  if (p.m31_slots + 4 * p.qm31_slots > BC_SLOTTED_TRACE_MAX_WORDS) p = lower(ops, true);
  return p;
}

}  // namespace vx

namespace vx {

// Straight-line HIP source of a program for one row per lane: every value is
// a local variable (the compiler allocates VGPRs), trace loads are read once
// at the top, constants and the row-offset steps are literals. Returns the
// kernel source; *n_ops = the number of instructions.
std::string bc_jit_source(const uint32_t* w, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                          uint32_t trace_log, size_t* n_ops) {
  const std::vector<Dec> ops = decode(w, n_words, n_trace_cols);
  if (n_ops) *n_ops = ops.size();
  std::string s;
  s.reserve(64 * ops.size() + 1024);
  s += "typedef unsigned int uint32_t;\ntypedef unsigned long long uint64_t;\ntypedef int int32_t;\n"
       "typedef long long int64_t;\n#include \"vx_field.h\"\nusing namespace vx;\n"
       "__device__ __forceinline__ qm31 add_base(qm31 x, uint32_t b) {\n"
       "  x.v[0] = m31_add(x.v[0], b);\n  return x;\n}\n"
       "extern \"C\" __global__ __launch_bounds__(256) void bc_jit(const uint32_t* const* __restrict__ cols, "
       "const uint32_t* __restrict__ coeff, const uint32_t* __restrict__ dinv, uint32_t* a0, uint32_t* a1, "
       "uint32_t* a2, uint32_t* a3, uint32_t n_rows) {\n"
       "  const uint32_t row = blockIdx.x * 256u + threadIdx.x;\n  if (row >= n_rows) return;\n"
       "  qm31 acc = {{0u, 0u, 0u, 0u}};\n";
  const uint32_t hmask = eval_log ? (1u << (eval_log - 1)) - 1u : 0u;
  bool offsets = false;
  for (const Dec& d : ops) offsets = offsets || (d.op == BC_LOAD_TRACE && d.offset != 0);
  if (offsets) {
    s += "  const uint32_t nat = __builtin_bitreverse32(row) >> " + std::to_string(32 - eval_log) + "u;\n";
    s += "  const uint32_t second = nat & ~" + std::to_string(hmask) + "u;\n";
  }
  auto qlit = [](const uint32_t* k) {
    return "qm31{{" + std::to_string(k[0]) + "u, " + std::to_string(k[1]) + "u, " + std::to_string(k[2]) + "u, " +
           std::to_string(k[3]) + "u}}";
  };
  // register -> (value name, is QM31)
  std::vector<std::pair<std::string, bool>> reg(1u << 16);
  std::vector<uint8_t> defined(1u << 16, 0);
  std::map<std::pair<uint32_t, int32_t>, std::string> trace;
  std::map<uint32_t, std::string> colp;
  uint32_t nv = 0, k = 0;
  auto m = [&](uint32_t r) {  // operand as M31
    if (!defined[r]) throw Error(ERR_ARG, "bytecode: register " + std::to_string(r) + " read before it is written");
    return reg[r].second ? reg[r].first + ".v[0]" : reg[r].first;
  };
  auto q = [&](uint32_t r) {  // operand as QM31
    if (!defined[r]) throw Error(ERR_ARG, "bytecode: register " + std::to_string(r) + " read before it is written");
    return reg[r].second ? reg[r].first : "qm31{{" + reg[r].first + ", 0u, 0u, 0u}}";
  };
  for (const Dec& d : ops) {
    std::string name = "v" + std::to_string(nv++), expr;
    bool wide = false;
    switch (d.op) {
      case BC_LOAD_CONST: expr = std::to_string(d.imm) + "u"; break;
      case BC_LOAD_SECURE_CONST: expr = qlit(d.k); wide = true; break;
      case BC_LOAD_TRACE: {
        auto key = std::make_pair(d.imm, d.offset);
        auto it = trace.find(key);
        if (it == trace.end()) {
          if (!colp.count(d.imm)) {
            colp[d.imm] = "c" + std::to_string(d.imm);
            s += "  const uint32_t* " + colp[d.imm] + " = cols[" + std::to_string(d.imm) + "];\n";
          }
          std::string rr = "row";
          if (d.offset != 0) {
            const int64_t half = (int64_t)1 << (eval_log - 1);
            int64_t st = ((int64_t)d.offset * ((int64_t)1 << (eval_log - trace_log - 1))) % half;
            if (st < 0) st += half;
            const std::string S = std::to_string((uint32_t)st) + "u", H = std::to_string(hmask) + "u";
            rr = "(__builtin_bitreverse32(second ? ((nat - " + S + ") & " + H + ") | second : (nat + " + S +
                 ") & " + H + ") >> " + std::to_string(32 - eval_log) + "u)";
          }
          const std::string t = "t" + std::to_string(trace.size());
          s += "  const uint32_t " + t + " = " + colp[d.imm] + " ? " + colp[d.imm] + "[" + rr + "] : 0u;\n";
          it = trace.emplace(key, t).first;
        }
        reg[d.dst] = {it->second, false};
        defined[d.dst] = 1;
        continue;
      }
      case BC_ADD: expr = "m31_add(" + m(d.src[0]) + ", " + m(d.src[1]) + ")"; break;
      case BC_SUB: expr = "m31_sub(" + m(d.src[0]) + ", " + m(d.src[1]) + ")"; break;
      case BC_MUL: expr = "m31_mul(" + m(d.src[0]) + ", " + m(d.src[1]) + ")"; break;
      case BC_NEG: expr = "m31_neg(" + m(d.src[0]) + ")"; break;
      case BC_ADD_CONST: expr = "m31_add(" + m(d.src[0]) + ", " + std::to_string(d.imm) + "u)"; break;
      case BC_MUL_CONST: expr = "m31_mul(" + m(d.src[0]) + ", " + std::to_string(d.imm) + "u)"; break;
      case BC_WIDE_ADD: expr = "qm_add(" + q(d.src[0]) + ", " + q(d.src[1]) + ")"; wide = true; break;
      case BC_WIDE_SUB: expr = "qm_sub(" + q(d.src[0]) + ", " + q(d.src[1]) + ")"; wide = true; break;
      case BC_WIDE_MUL: expr = "qm_mul(" + q(d.src[0]) + ", " + q(d.src[1]) + ")"; wide = true; break;
      case BC_WIDE_NEG: expr = "qm_sub(qm31{{0u, 0u, 0u, 0u}}, " + q(d.src[0]) + ")"; wide = true; break;
      case BC_WIDE_ADD_CONST: expr = "qm_add(" + q(d.src[0]) + ", " + qlit(d.k) + ")"; wide = true; break;
      case BC_WIDE_MUL_CONST: expr = "qm_mul(" + q(d.src[0]) + ", " + qlit(d.k) + ")"; wide = true; break;
      case BC_WIDE_ADD_BASE: expr = "add_base(" + q(d.src[0]) + ", " + m(d.src[1]) + ")"; wide = true; break;
      case BC_WIDE_MUL_BASE: expr = "qm_scale(" + q(d.src[0]) + ", " + m(d.src[1]) + ")"; wide = true; break;
      case BC_BASE_ADD_SECURE_CONST:
        expr = "qm_add(qm31{{" + m(d.src[0]) + ", 0u, 0u, 0u}}, " + qlit(d.k) + ")";
        wide = true;
        break;
      case BC_BASE_MUL_SECURE_CONST: expr = "qm_scale(" + qlit(d.k) + ", " + m(d.src[0]) + ")"; wide = true; break;
      case BC_WIDEN: expr = "qm31{{" + m(d.src[0]) + ", 0u, 0u, 0u}}"; wide = true; break;
      case BC_COMBINE_EF:
        expr = "qm31{{" + m(d.src[0]) + ", " + m(d.src[1]) + ", " + m(d.src[2]) + ", " + m(d.src[3]) + "}}";
        wide = true;
        break;
      case BC_ADD_CONSTRAINT: {
        const std::string kk = std::to_string(4 * k++);
        s += "  acc = qm_add(acc, qm_mul(" + q(d.src[0]) + ", qm31{{coeff[" + kk + "], coeff[" + kk + " + 1], coeff[" +
             kk + " + 2], coeff[" + kk + " + 3]}}));\n";
        continue;
      }
      default: break;
    }
    s += std::string("  const ") + (wide ? "qm31 " : "uint32_t ") + name + " = " + expr + ";\n";
    reg[d.dst] = {name, wide};
    defined[d.dst] = 1;
  }
  s += "  const qm31 res = qm_scale(acc, dinv[row >> " + std::to_string(trace_log) + "u]);\n"
       "  a0[row] = m31_add(a0[row], res.v[0]);\n  a1[row] = m31_add(a1[row], res.v[1]);\n"
       "  a2[row] = m31_add(a2[row], res.v[2]);\n  a3[row] = m31_add(a3[row], res.v[3]);\n}\n";
  return s;
}

}  // namespace vx
