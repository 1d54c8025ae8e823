// Batched uneven-column commitment: many M31 columns of different log sizes
// are interpolated on their trace domains, evaluated on their 2^blowup-times
// larger eval domains, bit-reversed, and committed into one lifted Blake2s
// Merkle tree (the reference's stwo-backend leaf rule,
// cuda/merkle_leaves_lifted.cu:1-98, columns grouped by ascending size). The
// reference batches only equal-size columns, one launch per layer and size
// (src/ntt.rs:191-281); here every phase is one launch driven by a device
// work list, whatever the number and sizes of the columns.
#pragma once
#include <memory>
#include <vector>

#include "prover.h"

namespace vx {

class BatchCommitter {
 public:
  // log_sizes sorted ascending, each in [14, 24]; log size + blowup <= 31.
  BatchCommitter(const uint32_t* log_sizes, uint32_t n_cols, uint32_t blowup_bits, bool bit_reverse,
                 hipStream_t stream);
  ~BatchCommitter();
  // Device column c: 0 = input values (2^L, kept intact), 1 = coefficients,
  // 2 = evaluation (2^(L+blowup), bit-reversed when enabled).
  uint32_t* column(uint32_t c, int which) const;
  uint32_t log_size(uint32_t c) const { return logs_[c]; }
  uint32_t n_cols() const { return (uint32_t)logs_.size(); }
  uint32_t lift_log() const { return lift_log_; }
  uint32_t blowup() const { return blowup_; }
  // LDE + bit reverse + lifted commitment, asynchronous on the stream; the
  // root lands in device memory (d_root()).
  void enqueue(bool timed);
  void wait(uint32_t root[8]);
  const uint32_t* d_root() const { return root_.u32(); }
  // ms of the last timed run: lde, bit reverse, commit, total
  void stage_times(float out[4]);
  uint64_t device_bytes() const { return device_bytes_; }
  hipStream_t stream() const { return stream_; }

 private:
  std::vector<uint32_t> logs_;
  uint32_t blowup_, lift_log_;
  bool bitrev_;
  hipStream_t stream_;
  bool own_stream_ = false;
  uint64_t device_bytes_ = 0;
  std::vector<std::shared_ptr<const ProverTables>> tables_;
  std::vector<uint64_t> off_in_, off_ev_;
  DeviceBuf in_, coef_, eval_, evr_, jobs_, br_jobs_, lift_, tree_, root_;
  LdeBatchStages st_{};
  uint32_t br_count_ = 0;
  uint64_t br_tiles_ = 0;
  TreeLevels tl_{};
  hipEvent_t ev_[4] = {};
  bool timed_ = false;
};

}  // namespace vx
