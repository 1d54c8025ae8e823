// Host pass of the bytecode constraint evaluator: decode the reference's
// register bytecode (crates/cuda-backend/src/constraint_eval/bytecode.rs:
// 173-323 writes it, cuda/constraint_eval.cu:122-407 reads it), check it,
// and rename its registers to typed slots for the device.
//
// The reference kernel gives every row thread a 1,024-entry QM31 register
// file in local memory. Here the program is straight-line code, so the live
// range of every value is known before the launch: a value lives from the
// instruction that writes it to its last read. A linear scan assigns each
// value a slot of its type (M31: one word per row, QM31: four), reusing
// slots whose value has died, so a wavefront's register file is as small as
// the program's peak number of live values and fits in LDS.
#include <algorithm>
#include <array>
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

BcProgram bc_compile(const uint32_t* w, uint32_t n_words, uint32_t n_trace_cols) {
  require(w != nullptr || n_words == 0, "bytecode: null program");
  // 1. decode
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

  // 2. values: every write of a register starts a new value; find each
  // value's type and last read
  const size_t n = ops.size();
  std::vector<int32_t> cur(1u << 16, -1);  // register -> live value
  std::vector<uint8_t> vtype;
  std::vector<size_t> last;
  std::vector<std::array<int32_t, 4>> src_val(n);
  std::vector<int32_t> dst_val(n, -1);
  for (size_t i = 0; i < n; i++) {
    const Dec& d = ops[i];
    for (uint32_t s = 0; s < d.n_src; s++) {
      const int32_t v = cur[d.src[s]];
      if (v < 0)
        throw Error(ERR_ARG, "bytecode: register " + std::to_string(d.src[s]) + " read before it is written");
      src_val[i][s] = v;
      last[v] = i;
    }
    if (d.op != BC_ADD_CONSTRAINT) {
      const int32_t v = (int32_t)vtype.size();
      vtype.push_back(result_type(d.op));
      last.push_back(i);
      cur[d.dst] = v;
      dst_val[i] = v;
    }
  }

  // 3. linear scan: free the slots of values read for the last time here,
  // then give the result a free slot of its type (the device reads every
  // operand before it writes the result, so the two may share a slot)
  BcProgram p;
  std::vector<uint32_t> slot(vtype.size(), 0);
  std::set<uint32_t> free_slots[2];
  uint32_t n_slots[2] = {0, 0};
  for (size_t i = 0; i < n; i++) {
    const Dec& d = ops[i];
    BcIns in{};
    uint32_t s_slot[4] = {0, 0, 0, 0}, s_type[4] = {0, 0, 0, 0};
    for (uint32_t s = 0; s < d.n_src; s++) {
      const int32_t v = src_val[i][s];
      s_slot[s] = slot[v];
      s_type[s] = vtype[v];
    }
    for (uint32_t s = 0; s < d.n_src; s++) {
      const int32_t v = src_val[i][s];
      bool dup = false;  // the same value read twice is freed once
      for (uint32_t t = 0; t < s; t++) dup = dup || src_val[i][t] == v;
      if (last[v] == i && !dup) free_slots[vtype[v]].insert(slot[v]);
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
    in.x = (d.op << 24) | dslot | (s_type[0] ? BC_TA : 0u) | (s_type[1] ? BC_TB : 0u) |
           (s_type[2] ? BC_TC : 0u) | (s_type[3] ? BC_TD : 0u);
    in.y = (s_slot[0] << 16) | s_slot[1];
    switch (d.op) {
      case BC_LOAD_CONST:
      case BC_ADD_CONST:
      case BC_MUL_CONST:
      case BC_LOAD_TRACE:
        in.z = d.imm;
        break;
      case BC_ADD_CONSTRAINT:
        in.z = p.n_constraints++;
        break;
      default:
        if (d.has_k) {
          in.z = (uint32_t)(p.consts.size() / 4);
          for (uint32_t x : d.k) p.consts.push_back(x);
        }
    }
    if (d.op == BC_LOAD_TRACE) {
      in.w = (uint32_t)d.offset;
      p.uses_offsets = p.uses_offsets || d.offset != 0;
    }
    if (d.op == BC_COMBINE_EF) in.w = (s_slot[2] << 16) | s_slot[3];
    p.ins.push_back(in);
  }
  p.m31_slots = n_slots[0];
  p.qm31_slots = n_slots[1];
  return p;
}

}  // namespace vx
