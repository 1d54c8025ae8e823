// Proof server: continuous batching of proof requests on one GPU.
//
// The reference serves proofs from one worker thread that drains a queue and
// proves one request at a time (src/bin/dashboard.rs:66-96; the batch API
// ProverPipeline::prove_batch, src/prover.rs:2278-2319, is the same serial
// loop). Here a dispatcher thread owns N slots, one HIP stream each. A
// request is started on a free slot as soon as one frees up; its proof
// overlaps on the device with the proofs running in the other slots. Every
// slot keeps a resident prover per size class (tables shared between slots),
// created on first use and admitted against a device-memory budget; idle
// provers are evicted least-recently-used when a new class does not fit.
// With one slot the server is the reference's one-at-a-time loop.
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "prover.h"

namespace vx {

struct ProofRequest {
  uint32_t air;          // AIR_FIBONACCI or AIR_D2
  uint32_t log_n;
  uint32_t blowup_bits;
  uint32_t n_queries;
  uint32_t pow_bits;
  uint32_t a, b;         // Fibonacci public inputs
  uint32_t n_cols, n_cons;  // AIR_D2 shape
  uint64_t seed;         // AIR_D2 trace seed
  uint64_t air_seed;     // AIR_D2 constraint coefficients
};

// What the server needs of a resident prover of one size class.
class ServedProver {
 public:
  virtual ~ServedProver() = default;
  // Enqueue the proof of `r` on the prover's stream (asynchronous).
  virtual void start(const ProofRequest& r) = 0;
  // Host bytes of the last proof; valid once the stream has passed the
  // proof, and until the proof after the next one starts.
  virtual const uint8_t* result() const = 0;
  virtual size_t result_bytes() const = 0;
  virtual uint64_t device_bytes() const = 0;
};

struct JobTimes {
  double submit_s, start_s, done_s;  // seconds since the server's epoch
};

class ProofServer {
 public:
  ProofServer(uint32_t n_slots, uint64_t mem_budget_bytes, bool graphs);
  ~ProofServer();

  // Create the size class of `r` on every slot now (no proof is run), so the
  // first requests of that class do not pay for allocation and tables.
  void prepare(const ProofRequest& r);
  // Queue a request; the proof bytes are written to `out` (capacity `cap`,
  // at least the class's proof size) before the ticket completes.
  uint64_t submit(const ProofRequest& r, uint8_t* out, size_t cap);
  // Block until the ticket completes. Returns false and sets `err` if the
  // request failed. The ticket is released afterwards.
  bool wait(uint64_t ticket, size_t* len, JobTimes* t, std::string& err);
  // Non-blocking: true once the ticket has completed (successfully or not).
  bool done(uint64_t ticket);

  double now() const;  // seconds since the server's epoch
  uint64_t resident_bytes() const { return resident_.load(); }
  uint32_t n_slots() const { return (uint32_t)slots_.size(); }

 private:
  using Key = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint64_t>;
  struct Job {
    ProofRequest req;
    uint8_t* out;
    size_t cap, len = 0;
    JobTimes t{};
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    std::string err;
  };
  struct Resident {
    std::unique_ptr<ServedProver> prover;
    uint64_t bytes;
    uint64_t last_use;
  };
  struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::map<Key, Resident> provers;
    uint64_t ticket = 0;  // 0 = free
    ServedProver* active = nullptr;
  };

  static Key key_of(const ProofRequest& r);
  // The slot's prover for r's class, created (and admitted against the
  // budget) if needed; nullptr when it does not fit until a slot frees up.
  ServedProver* acquire(Slot& sl, const ProofRequest& r);
  std::unique_ptr<ServedProver> create(const ProofRequest& r, hipStream_t s, uint64_t* bytes);
  static uint64_t estimate(const ProofRequest& r);
  bool evict_one(uint64_t need);
  void loop();

  std::vector<Slot> slots_;
  uint64_t budget_;
  bool graphs_;
  std::atomic<uint64_t> resident_{0};
  uint64_t use_clock_ = 0;
  std::chrono::steady_clock::time_point epoch_;

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<uint64_t> incoming_;
  std::unordered_map<uint64_t, Job> jobs_;
  uint64_t next_ticket_ = 1;
  bool stop_ = false;
  std::thread thread_;
  std::mutex prov_mu_;  // slot prover maps and tables (dispatcher vs prepare)
  std::map<Key, uint64_t> table_bytes_;  // shared tables, counted once per class
  int device_ = 0;
};

}  // namespace vx
