// Scheduler of the proof server: slots, resident provers per size class,
// admission against a device-memory budget, eviction and proof copy-out.
//
// The reference serves proofs from one worker thread that drains a queue and
// proves one request at a time (src/bin/dashboard.rs:66-96; the batch API
// ProverPipeline::prove_batch, src/prover.rs:2278-2319, is the same serial
// loop). Here a dispatcher thread owns N slots. A request is started on a
// free slot as soon as one frees up, so proofs of different slots overlap.
// Every slot keeps a resident prover per size class, created on first use
// and admitted against a device-memory budget; idle provers are evicted
// least-recently-used when a new class does not fit. With one slot the
// server is the reference's one-at-a-time loop.
//
// This part holds no HIP code: provers come from a factory (the HIP provers
// in server.cpp, mock provers in the host-only scheduler tests), so every
// interleaving of completion, eviction and copy-out can be forced on a CPU.
#pragma once
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

namespace vx {

struct ProofRequest {
  uint32_t air;          // 0 = Fibonacci, 1 = synthetic multi-column AIR
  uint32_t log_n;
  uint32_t blowup_bits;
  uint32_t n_queries;
  uint32_t pow_bits;
  uint32_t a, b;         // Fibonacci public inputs
  uint32_t n_cols, n_cons;  // multi-column AIR shape
  uint64_t seed;         // multi-column AIR trace seed
  uint64_t air_seed;     // multi-column AIR constraint coefficients
};

// What the scheduler needs of a resident prover of one size class.
class ServedProver {
 public:
  virtual ~ServedProver() = default;
  // Enqueue the proof of `r` (asynchronous); throws if it cannot be started.
  virtual void start(const ProofRequest& r) = 0;
  // State of the last started proof: 0 running, 1 finished, -1 failed
  // (reason in *err).
  virtual int poll(std::string* err) = 0;
  // Bytes of the last finished proof. They stay valid while at most one
  // further proof is started on this prover (two result buffers alternate),
  // and until the prover is destroyed.
  virtual const uint8_t* result() const = 0;
  virtual size_t result_bytes() const = 0;
  virtual uint64_t device_bytes() const = 0;
};

struct ServerHooks {
  // The prover of r's class for slot `slot` (throws on failure).
  std::function<std::unique_ptr<ServedProver>(uint32_t slot, const ProofRequest& r)> create;
  // Admission estimates: device bytes of one prover of r's class (shared
  // tables excluded), and of the tables every prover of r's (log_n, blowup)
  // shares on the device.
  std::function<uint64_t(const ProofRequest& r)> prover_bytes;
  std::function<uint64_t(const ProofRequest& r)> table_bytes;
  // Run once on the dispatcher thread before it serves (e.g. select the
  // device); may be empty.
  std::function<void()> thread_init;
};

struct JobTimes {
  double submit_s, start_s, done_s;  // seconds since the server's epoch
};

struct ServerStats {
  uint64_t created;   // provers built
  uint64_t evicted;   // provers destroyed to make room
  uint64_t deferred;  // admissions postponed because the only candidates were
                      // running or still held an uncopied proof
};

class ServerCore {
 public:
  ServerCore(uint32_t n_slots, uint64_t mem_budget_bytes, ServerHooks hooks);
  // Finishes every queued request, then stops the dispatcher and destroys
  // the provers.
  ~ServerCore();

  // Create the class of `r` on every slot now (no proof is run); fails if
  // that does not fit the budget without evicting.
  void prepare(const ProofRequest& r);
  // Queue a request; the proof bytes are written to `out` (capacity `cap`)
  // before the ticket completes.
  uint64_t submit(const ProofRequest& r, uint8_t* out, size_t cap);
  // Block until the ticket completes. Returns false and sets `err` if the
  // request failed. The ticket is released afterwards.
  bool wait(uint64_t ticket, size_t* len, JobTimes* t, std::string& err);
  // Non-blocking: true once the ticket has completed (successfully or not).
  bool done(uint64_t ticket);

  double now() const;  // seconds since the server's epoch
  uint64_t resident_bytes() const { return resident_.load(); }
  uint32_t n_slots() const { return (uint32_t)slots_.size(); }
  ServerStats stats() const;
  // Device bytes the budget is charged for, recomputed from the resident
  // provers and the shapes they use (equals resident_bytes()).
  uint64_t recount_bytes();

  using Key = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint64_t>;
  using Shape = std::pair<uint32_t, uint32_t>;  // (log_n, blowup_bits): what the tables depend on
  static Key key_of(const ProofRequest& r);
  static Shape shape_of(const ProofRequest& r) { return {r.log_n, r.blowup_bits}; }

 private:
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
    Shape shape;
  };
  struct TableUse {
    uint64_t bytes = 0;  // charged while refs > 0
    uint32_t refs = 0;   // resident provers of this shape
  };
  // Dispatcher-thread state of a slot. `provers` is guarded by prov_mu_;
  // the other fields are only touched by the dispatcher thread.
  struct Slot {
    std::map<Key, Resident> provers;
    uint64_t ticket = 0;  // running request, 0 = free
    ServedProver* active = nullptr;
    Key active_key{};
    bool pending = false;  // a finished proof of `pending_key` awaits copy-out
    Key pending_key{};
  };

  // The slot's prover for r's class, created (and admitted against the
  // budget) if needed; nullptr when it has to wait for a running proof or a
  // copy-out to release memory. Throws when it can never fit.
  ServedProver* acquire(uint32_t slot, const ProofRequest& r);
  bool evictable(const Slot& sl, const Key& k) const;
  bool evict_one();  // caller holds prov_mu_
  void admit(uint32_t slot, const Key& k, const ProofRequest& r);  // caller holds prov_mu_
  uint64_t admission_need(const ProofRequest& r) const;  // caller holds prov_mu_
  void loop();

  ServerHooks hooks_;
  std::vector<Slot> slots_;
  uint64_t budget_;
  std::atomic<uint64_t> resident_{0};
  uint64_t use_clock_ = 0;
  std::map<Shape, TableUse> tables_;
  ServerStats stats_{};
  std::chrono::steady_clock::time_point epoch_;

  std::mutex mu_;  // jobs, queue, stop flag
  std::condition_variable cv_work_, cv_done_;
  std::deque<uint64_t> incoming_;
  std::unordered_map<uint64_t, Job> jobs_;
  uint64_t next_ticket_ = 1;
  bool stop_ = false;
  mutable std::mutex prov_mu_;  // slot prover maps, tables_, stats_, use_clock_
  std::thread thread_;
};

}  // namespace vx
