// Continuous-batching proof server (see server.h).
#include "server.h"

#include <algorithm>
#include <cstring>

#include "vx_d2.h"

namespace vx {

namespace {

AirSpec air_of(const ProofRequest& r) {
  AirSpec a;
  a.kind = r.air;
  if (r.air == AIR_D2) {
    a.n_cols = r.n_cols;
    a.n_cons = r.n_cons;
    a.air_seed = r.air_seed;
  }
  return a;
}

class Served final : public ServedProver {
 public:
  Served(const ProofRequest& r, hipStream_t s, bool graphs)
      : p_(r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, s, air_of(r)) {
    p_.set_graphs(graphs);
    if (graphs) p_.prepare_graph();
  }
  void start(const ProofRequest& r) override {
    if (r.air == AIR_D2) p_.enqueue((uint32_t)r.seed, (uint32_t)(r.seed >> 32), false);
    else p_.enqueue(r.a, r.b, false);
  }
  const uint8_t* result() const override { return p_.host_proof(); }
  size_t result_bytes() const override { return p_.proof_bytes(); }
  uint64_t device_bytes() const override { return p_.device_bytes(); }

 private:
  StarkProver p_;
};

void check_request(const ProofRequest& r) {
  require(r.air == AIR_FIBONACCI || r.air == AIR_D2, "server: unknown AIR");
  require(r.log_n >= 4 && r.blowup_bits >= 1 && r.log_n + r.blowup_bits <= 31,
          "server: need log_n >= 4, blowup >= 1, log_n + blowup <= 31");
  require(r.n_queries >= 1 && r.n_queries <= 65536 && r.pow_bits <= 32, "server: bad n_queries / pow_bits");
  if (r.air == AIR_FIBONACCI) require(r.a < P && r.b < P, "server: public inputs must be canonical M31");
  else require(r.n_cols >= 1 && r.n_cols <= D2_MAX_COLS && r.n_cons >= 1 && r.n_cons <= D2_MAX_CONS,
               "server: multi-column AIR needs 1..256 columns and 1..1024 constraints");
}

}  // namespace

ProofServer::ProofServer(uint32_t n_slots, uint64_t mem_budget_bytes, bool graphs)
    : budget_(mem_budget_bytes), graphs_(graphs), epoch_(std::chrono::steady_clock::now()) {
  require(n_slots >= 1 && n_slots <= 64, "server: n_slots must be 1..64");
  VX_HIP(hipGetDevice(&device_));
  slots_.resize(n_slots);
  for (auto& sl : slots_) {
    VX_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    VX_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  }
  thread_ = std::thread([this] { loop(); });
}

ProofServer::~ProofServer() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (thread_.joinable()) thread_.join();
  for (auto& sl : slots_) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    sl.provers.clear();
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
}

double ProofServer::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch_).count();
}

ProofServer::Key ProofServer::key_of(const ProofRequest& r) {
  const bool d2 = r.air == AIR_D2;
  return Key{r.air, r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, d2 ? r.n_cols : 0, d2 ? r.n_cons : 0,
             d2 ? r.air_seed : 0};
}

uint64_t ProofServer::estimate(const ProofRequest& r) {
  return StarkProver::estimate_bytes(r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, air_of(r));
}

std::unique_ptr<ServedProver> ProofServer::create(const ProofRequest& r, hipStream_t s, uint64_t* bytes) {
  auto p = std::make_unique<Served>(r, s, graphs_);
  *bytes = p->device_bytes();
  return p;
}

// Evict the least recently used prover of an idle slot (freeing it then
// waits for nothing); false if there is none.
bool ProofServer::evict_one(uint64_t) {
  Slot* best_slot = nullptr;
  Key best_key{};
  uint64_t best_use = ~0ull;
  for (auto& sl : slots_)
    if (!sl.ticket)
      for (auto& [k, res] : sl.provers)
        if (res.last_use < best_use) {
        best_use = res.last_use;
        best_slot = &sl;
        best_key = k;
      }
  if (!best_slot) return false;
  auto it = best_slot->provers.find(best_key);
  resident_ -= it->second.bytes;
  best_slot->provers.erase(it);
  return true;
}

ServedProver* ProofServer::acquire(Slot& sl, const ProofRequest& r) {
  std::lock_guard<std::mutex> lk(prov_mu_);
  const Key k = key_of(r);
  auto it = sl.provers.find(k);
  if (it != sl.provers.end()) {
    it->second.last_use = ++use_clock_;
    return it->second.prover.get();
  }
  uint64_t tables = table_bytes_.count(k) ? 0 : ProverTables::estimate_bytes(r.log_n, r.blowup_bits);
  const uint64_t need = estimate(r) + tables;
  while (resident_ + need > budget_) {
    if (!evict_one(need)) {
      bool busy = std::any_of(slots_.begin(), slots_.end(), [](const Slot& x) { return x.ticket != 0; });
      if (busy) return nullptr;  // retry once a running proof releases its prover
      throw Error(ERR_ALLOC, "server: request does not fit the device-memory budget");
    }
  }
  uint64_t bytes = 0;
  auto p = create(r, sl.stream, &bytes);
  if (tables) {
    table_bytes_[k] = tables;
    resident_ += tables;
  }
  resident_ += bytes;
  ServedProver* raw = p.get();
  sl.provers.emplace(k, Resident{std::move(p), bytes, ++use_clock_});
  return raw;
}

void ProofServer::prepare(const ProofRequest& r) {
  check_request(r);
  // Serialise with the dispatcher: it never runs a proof of a class that
  // is being created (the class is new), and the slot maps are locked.
  VX_HIP(hipSetDevice(device_));
  for (auto& sl : slots_) {
    std::lock_guard<std::mutex> lk(prov_mu_);
    const Key k = key_of(r);
    if (sl.provers.count(k)) continue;
    uint64_t tables = table_bytes_.count(k) ? 0 : ProverTables::estimate_bytes(r.log_n, r.blowup_bits);
    require(resident_ + estimate(r) + tables <= budget_, "server: prepare exceeds the device-memory budget");
    uint64_t bytes = 0;
    auto p = create(r, sl.stream, &bytes);
    if (tables) {
      table_bytes_[k] = tables;
      resident_ += tables;
    }
    resident_ += bytes;
    sl.provers.emplace(k, Resident{std::move(p), bytes, ++use_clock_});
  }
}

uint64_t ProofServer::submit(const ProofRequest& r, uint8_t* out, size_t cap) {
  check_request(r);
  require(out != nullptr, "server: null output buffer");
  const size_t need = proof_size(r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, air_of(r));
  require(cap >= need, "server: output buffer smaller than the proof (see vx_proof_size)");
  uint64_t t;
  {
    std::lock_guard<std::mutex> lk(mu_);
    require(!stop_, "server: stopped");
    t = next_ticket_++;
    Job& j = jobs_[t];
    j.req = r;
    j.out = out;
    j.cap = cap;
    j.t.submit_s = now();
    incoming_.push_back(t);
  }
  cv_work_.notify_one();
  return t;
}

bool ProofServer::done(uint64_t ticket) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(ticket);
  require(it != jobs_.end(), "server: unknown ticket");
  return it->second.state >= 2;
}

bool ProofServer::wait(uint64_t ticket, size_t* len, JobTimes* t, std::string& err) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = jobs_.find(ticket);
  require(it != jobs_.end(), "server: unknown ticket");
  Job* j = &it->second;  // element references survive rehashing, iterators do not
  cv_done_.wait(lk, [&] { return j->state >= 2; });
  const bool ok = j->state == 2;
  if (len) *len = j->len;
  if (t) *t = j->t;
  err = j->err;
  jobs_.erase(ticket);
  return ok;
}

// Dispatcher: polls the slots' completion events, starts queued requests on
// free slots, then copies finished proofs out of the provers' pinned buffers
// (after the slot has already been refilled, so the GPU does not wait for it).
void ProofServer::loop() {
  (void)hipSetDevice(device_);
  std::deque<uint64_t> queue;
  struct Finished {
    uint64_t ticket;
    const uint8_t* src;
    size_t len;
    double t;
    std::string err;
  };
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      auto busy = [&] {
        return std::any_of(slots_.begin(), slots_.end(), [](const Slot& x) { return x.ticket != 0; });
      };
      if (queue.empty() && incoming_.empty() && !busy()) {
        if (stop_) return;
        cv_work_.wait(lk, [&] { return stop_ || !incoming_.empty(); });
        if (stop_ && incoming_.empty()) return;
      }
      while (!incoming_.empty()) {
        queue.push_back(incoming_.front());
        incoming_.pop_front();
      }
    }
    bool progress = false;
    std::vector<Finished> fin;
    for (auto& sl : slots_) {
      if (!sl.ticket) continue;
      hipError_t e = hipEventQuery(sl.done);
      if (e == hipErrorNotReady) continue;
      Finished f{sl.ticket, nullptr, 0, now(), ""};
      if (e == hipSuccess) {
        f.src = sl.active->result();
        f.len = sl.active->result_bytes();
      } else {
        f.err = std::string("server: proof failed on the device: ") + hipGetErrorString(e);
      }
      fin.push_back(std::move(f));
      sl.ticket = 0;
      sl.active = nullptr;
      progress = true;
    }
    while (!queue.empty()) {
      Slot* sl = nullptr;
      const uint64_t tk = queue.front();
      ProofRequest r;
      {
        std::lock_guard<std::mutex> lk(mu_);
        r = jobs_[tk].req;
      }
      {  // prefer a free slot that already holds this class
        std::lock_guard<std::mutex> lk(prov_mu_);
        for (auto& x : slots_)
          if (!x.ticket && x.provers.count(key_of(r))) {
            sl = &x;
            break;
          }
      }
      if (!sl)
        for (auto& x : slots_)
          if (!x.ticket) {
            sl = &x;
            break;
          }
      if (!sl) break;
      std::string err;
      try {
        ServedProver* p = acquire(*sl, r);
        if (!p) break;  // waits for memory
        {
          std::lock_guard<std::mutex> lk(mu_);
          Job& j = jobs_[tk];
          j.state = 1;
          j.t.start_s = now();
        }
        p->start(r);
        VX_HIP(hipEventRecord(sl->done, sl->stream));
        sl->ticket = tk;
        sl->active = p;
      } catch (const std::exception& ex) {
        err = ex.what();
      }
      if (!err.empty()) fin.push_back(Finished{tk, nullptr, 0, now(), err});
      queue.pop_front();
      progress = true;
    }
    if (!fin.empty()) {
      for (auto& f : fin) {
        Job* j;
        {
          std::lock_guard<std::mutex> lk(mu_);
          j = &jobs_[f.ticket];
        }
        if (f.err.empty()) std::memcpy(j->out, f.src, f.len);
        std::lock_guard<std::mutex> lk(mu_);
        j->len = f.len;
        j->t.done_s = f.t;
        j->err = f.err;
        j->state = f.err.empty() ? 2 : 3;
      }
      cv_done_.notify_all();
    }
    if (!progress) std::this_thread::yield();
  }
}

}  // namespace vx
