// Proof server on one GPU: the scheduler of server_core.h over HIP provers.
// Each slot owns one HIP stream (default priority); a slot's provers (one per
// size class) enqueue their proofs on it, so proofs of different slots
// overlap on the device. Twiddle tables are shared by every prover of a
// (log_n, blowup) shape on the device.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "prover.h"
#include "server_core.h"

namespace vx {

class ProofServer {
 public:
  ProofServer(uint32_t n_slots, uint64_t mem_budget_bytes, bool graphs);
  // Finishes queued requests, destroys the provers, then the slot streams.
  ~ProofServer();

  void prepare(const ProofRequest& r);
  uint64_t submit(const ProofRequest& r, uint8_t* out, size_t cap);
  bool wait(uint64_t ticket, size_t* len, JobTimes* t, std::string& err) { return core_->wait(ticket, len, t, err); }
  bool done(uint64_t ticket) { return core_->done(ticket); }
  double now() const { return core_->now(); }
  uint64_t resident_bytes() const { return core_->resident_bytes(); }
  ServerStats stats() const { return core_->stats(); }
  uint32_t n_slots() const { return core_->n_slots(); }

 private:
  int device_ = 0;
  bool graphs_;
  std::vector<hipStream_t> streams_;
  std::unique_ptr<ServerCore> core_;
};

}  // namespace vx
