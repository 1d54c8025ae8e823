// Generic constraint evaluation from the stwo backend's register bytecode
// (cuda/constraint_eval.cu, crates/cuda-backend/src/constraint_eval/
// bytecode.rs): opcodes, the compact instruction form the host pass writes
// for the device, and the circle-domain row offset. Shared by the host pass,
// the kernel and the tests.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "vx_field.h"

namespace vx {

// Opcodes of the reference encoding (cuda/constraint_eval.cu:19-43).
enum BcOp : uint32_t {
  BC_LOAD_CONST = 0x01,
  BC_LOAD_SECURE_CONST = 0x02,
  BC_LOAD_TRACE = 0x03,
  BC_ADD = 0x10,
  BC_SUB = 0x11,
  BC_MUL = 0x12,
  BC_NEG = 0x13,
  BC_ADD_CONST = 0x14,
  BC_MUL_CONST = 0x15,
  BC_WIDE_ADD = 0x20,
  BC_WIDE_SUB = 0x21,
  BC_WIDE_MUL = 0x22,
  BC_WIDE_NEG = 0x23,
  BC_WIDE_ADD_CONST = 0x24,
  BC_WIDE_MUL_CONST = 0x25,
  BC_WIDE_ADD_BASE = 0x28,
  BC_WIDE_MUL_BASE = 0x29,
  BC_BASE_ADD_SECURE_CONST = 0x2A,
  BC_BASE_MUL_SECURE_CONST = 0x2B,
  BC_WIDEN = 0x30,
  BC_COMBINE_EF = 0x31,
  BC_ADD_CONSTRAINT = 0x40,
};

// Compact instruction (16 bytes, one scalar load on the device). Values are
// renamed to typed slots (an M31 slot holds one word per row, a QM31 slot
// four). Trace loads take no instruction: BcProgram::trace lists each
// distinct (column, offset) once with its M31 slot, filled by a prologue, or
// (remat) with no slot: then an operand of kind TRACE names the entry and
// reads the column where it is used.
// Constants take no slot: an operand of kind CONST is a constant-pool entry
// (M31 constants as (v, 0, 0, 0)). An operand read as QM31 from an M31
// source is (v, 0, 0, 0); read as M31 from a QM31 source it is v[0].
//   x: op (bits 24..31) | kinds of operands a, b, c, d (2 bits each from
//      bit 16) | dst slot (bits 0..15)
//   y: a (bits 16..31) | b (bits 0..15)   slot or constant index
//   z: M31 constant (ADD_CONST, MUL_CONST), constant index (QM31 constant
//      operations) or constraint index
//   w: c << 16 | d (COMBINE_EF)
struct BcIns {
  uint32_t x, y, z, w;
};
enum : uint32_t { BC_K_M = 0, BC_K_Q = 1, BC_K_TRACE = 2, BC_K_CONST = 3 };
constexpr uint32_t BC_NO_SLOT = 0xFFFFFFFFu;
// Largest register file (words per row) for which trace values take slots.
constexpr uint32_t BC_SLOTTED_TRACE_MAX_WORDS = 32;
constexpr uint32_t BC_KA = 16, BC_KB = 18, BC_KC = 20, BC_KD = 22;

struct BcProgram {
  std::vector<BcIns> ins;
  std::vector<uint32_t> consts;  // QM31 constants, 4 words each
  std::vector<uint32_t> trace;   // (column, offset, M31 slot) of each distinct trace load
  uint32_t m31_slots = 0, qm31_slots = 0, n_constraints = 0;
  bool uses_offsets = false;
  bool remat = false;
};

// Decodes and checks the reference bytecode (n_words words) for columns
// [0, n_trace_cols), renames registers to the fewest slots of each type
// (a linear scan over the straight-line program: a slot is reused once the
// value in it has had its last read). Throws vx::Error on malformed
// programs, unknown opcodes, a read of a register never written, or a
// column index out of range.
BcProgram bc_compile(const uint32_t* words, uint32_t n_words, uint32_t n_trace_cols);

// Straight-line HIP source of the program (kernel "bc_jit": one row per
// lane, 256 lanes per workgroup) for an evaluation of 2^eval_log rows over a
// trace domain of 2^trace_log; *n_ops = the program's instruction count.
std::string bc_jit_source(const uint32_t* words, uint32_t n_words, uint32_t n_trace_cols, uint32_t eval_log,
                          uint32_t trace_log, size_t* n_ops);

// stwo's offset_bit_reversed_circle_domain_index (cuda/constraint_eval.cu:
// 63-94): the bit-reversed row `offset` trace-domain steps from row i of an
// evaluation of 2^eval_log rows over a trace domain of 2^domain_log rows
// (eval_log > domain_log). The offset is taken modulo the half domain
// without 32-bit overflow.
VX_HD uint32_t bc_offset_row(uint32_t i, uint32_t domain_log, uint32_t eval_log, int32_t offset) {
  const uint32_t sh = 32u - eval_log;
  uint32_t prev = eval_log ? (__builtin_bitreverse32(i) >> sh) : 0u;
  const uint32_t half = 1u << (eval_log - 1);
  const int64_t step = (int64_t)offset << (eval_log - domain_log - 1);
  const int64_t h = (int64_t)half;
  if (prev < half) {
    int64_t idx = ((int64_t)prev + step) % h;
    if (idx < 0) idx += h;
    prev = (uint32_t)idx;
  } else {
    int64_t idx = ((int64_t)(prev - half) - step) % h;
    if (idx < 0) idx += h;
    prev = (uint32_t)idx + half;
  }
  return __builtin_bitreverse32(prev) >> sh;
}

}  // namespace vx
