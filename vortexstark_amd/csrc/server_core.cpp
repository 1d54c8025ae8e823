// Scheduler of the proof server (see server_core.h). Host code only.
#include "server_core.h"

#include <algorithm>
#include <cstring>

#include "vx_error.h"

namespace vx {

ServerCore::ServerCore(uint32_t n_slots, uint64_t mem_budget_bytes, ServerHooks hooks)
    : hooks_(std::move(hooks)), budget_(mem_budget_bytes), epoch_(std::chrono::steady_clock::now()) {
  require(n_slots >= 1 && n_slots <= 64, "server: n_slots must be 1..64");
  require(hooks_.create && hooks_.prover_bytes && hooks_.table_bytes, "server: incomplete prover hooks");
  slots_.resize(n_slots);
  thread_ = std::thread([this] { loop(); });
}

ServerCore::~ServerCore() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (thread_.joinable()) thread_.join();
  std::lock_guard<std::mutex> lk(prov_mu_);
  for (auto& sl : slots_) sl.provers.clear();
}

double ServerCore::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch_).count();
}

ServerStats ServerCore::stats() const {
  std::lock_guard<std::mutex> lk(prov_mu_);
  return stats_;
}

ServerCore::Key ServerCore::key_of(const ProofRequest& r) {
  const bool multi = r.air != 0;
  return Key{r.air,   r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, multi ? r.n_cols : 0, multi ? r.n_cons : 0,
             multi ? r.air_seed : 0};
}

uint64_t ServerCore::recount_bytes() {
  std::lock_guard<std::mutex> lk(prov_mu_);
  std::map<Shape, uint32_t> refs;
  uint64_t b = 0;
  for (auto& sl : slots_)
    for (auto& kv : sl.provers) {
      b += kv.second.bytes;
      refs[kv.second.shape]++;
    }
  for (auto& kv : tables_) {
    auto it = refs.find(kv.first);
    const uint32_t want = it == refs.end() ? 0 : it->second;
    if (kv.second.refs != want) throw Error(ERR_INTERNAL, "server: table reference count out of step");
    if (kv.second.refs) b += kv.second.bytes;
  }
  return b;
}

// A prover may be destroyed only on a slot with no proof running (its
// destructor waits for the slot's stream) and only once the proof it last
// finished has been copied out (result() points into its own buffers).
bool ServerCore::evictable(const Slot& sl, const Key& k) const {
  return sl.ticket == 0 && !(sl.pending && sl.pending_key == k);
}

bool ServerCore::evict_one() {
  Slot* best_slot = nullptr;
  Key best_key{};
  uint64_t best_use = ~0ull;
  for (auto& sl : slots_)
    for (auto& kv : sl.provers)
      if (evictable(sl, kv.first) && kv.second.last_use < best_use) {
        best_use = kv.second.last_use;
        best_slot = &sl;
        best_key = kv.first;
      }
  if (!best_slot) return false;
  auto it = best_slot->provers.find(best_key);
  resident_ -= it->second.bytes;
  TableUse& tu = tables_[it->second.shape];
  if (--tu.refs == 0) {
    resident_ -= tu.bytes;
    tu.bytes = 0;
  }
  best_slot->provers.erase(it);
  stats_.evicted++;
  return true;
}

uint64_t ServerCore::admission_need(const ProofRequest& r) const {
  auto it = tables_.find(shape_of(r));
  const bool shared = it != tables_.end() && it->second.refs > 0;
  return hooks_.prover_bytes(r) + (shared ? 0 : hooks_.table_bytes(r));
}

void ServerCore::admit(uint32_t slot, const Key& k, const ProofRequest& r) {
  std::unique_ptr<ServedProver> p = hooks_.create(slot, r);  // nothing is charged if this throws
  const uint64_t bytes = std::max(p->device_bytes(), hooks_.prover_bytes(r));
  TableUse& tu = tables_[shape_of(r)];
  if (tu.refs++ == 0) {
    tu.bytes = hooks_.table_bytes(r);
    resident_ += tu.bytes;
  }
  resident_ += bytes;
  slots_[slot].provers.emplace(k, Resident{std::move(p), bytes, ++use_clock_, shape_of(r)});
  stats_.created++;
}

ServedProver* ServerCore::acquire(uint32_t slot, const ProofRequest& r) {
  std::lock_guard<std::mutex> lk(prov_mu_);
  Slot& sl = slots_[slot];
  const Key k = key_of(r);
  auto it = sl.provers.find(k);
  if (it != sl.provers.end()) {
    it->second.last_use = ++use_clock_;
    return it->second.prover.get();
  }
  while (resident_ + admission_need(r) > budget_) {
    if (!evict_one()) {
      const bool waitable =
          std::any_of(slots_.begin(), slots_.end(), [](const Slot& x) { return x.ticket != 0 || x.pending; });
      if (waitable) {  // a running proof or a copy-out will release a prover
        stats_.deferred++;
        return nullptr;
      }
      throw Error(ERR_ALLOC, "server: request does not fit the device-memory budget");
    }
  }
  admit(slot, k, r);
  return sl.provers.find(k)->second.prover.get();
}

void ServerCore::prepare(const ProofRequest& r) {
  const Key k = key_of(r);
  for (uint32_t s = 0; s < slots_.size(); s++) {
    std::lock_guard<std::mutex> lk(prov_mu_);
    if (slots_[s].provers.count(k)) continue;
    require(resident_ + admission_need(r) <= budget_, "server: prepare exceeds the device-memory budget");
    admit(s, k, r);
  }
}

uint64_t ServerCore::submit(const ProofRequest& r, uint8_t* out, size_t cap) {
  require(out != nullptr, "server: null output buffer");
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

bool ServerCore::done(uint64_t ticket) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(ticket);
  require(it != jobs_.end(), "server: unknown ticket");
  return it->second.state >= 2;
}

bool ServerCore::wait(uint64_t ticket, size_t* len, JobTimes* t, std::string& err) {
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

// Dispatcher. Each pass: (1) poll the running slots; a finished proof's
// slot becomes free and its prover is marked as holding an uncopied result;
// (2) start queued requests on free slots (creating or evicting provers as
// needed; a marked prover is never evicted); (3) copy the finished proofs
// out and clear the marks. Copying after refilling keeps the device busy
// while the host copies.
void ServerCore::loop() {
  if (hooks_.thread_init) {
    try {
      hooks_.thread_init();
    } catch (...) {
    }
  }
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
      const bool busy = std::any_of(slots_.begin(), slots_.end(), [](const Slot& x) { return x.ticket != 0; });
      if (queue.empty() && incoming_.empty() && !busy) {
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
      std::string why;
      int st;
      try {
        st = sl.active->poll(&why);
      } catch (const std::exception& ex) {
        st = -1;
        why = ex.what();
      }
      if (st == 0) continue;
      Finished f{sl.ticket, nullptr, 0, now(), ""};
      if (st > 0) {
        f.src = sl.active->result();
        f.len = sl.active->result_bytes();
        sl.pending = true;
        sl.pending_key = sl.active_key;
      } else {
        f.err = "server: proof failed on the device: " + why;
      }
      fin.push_back(std::move(f));
      sl.ticket = 0;
      sl.active = nullptr;
      progress = true;
    }
    while (!queue.empty()) {
      const uint64_t tk = queue.front();
      ProofRequest r;
      {
        std::lock_guard<std::mutex> lk(mu_);
        r = jobs_[tk].req;
      }
      const Key k = key_of(r);
      int pick = -1;
      {  // prefer a free slot that already holds this class
        std::lock_guard<std::mutex> lk(prov_mu_);
        for (uint32_t s = 0; s < slots_.size() && pick < 0; s++)
          if (!slots_[s].ticket && slots_[s].provers.count(k)) pick = (int)s;
      }
      for (uint32_t s = 0; s < slots_.size() && pick < 0; s++)
        if (!slots_[s].ticket) pick = (int)s;
      if (pick < 0) break;
      Slot& sl = slots_[pick];
      std::string err;
      try {
        ServedProver* p = acquire((uint32_t)pick, r);
        if (!p) break;  // waits for a running proof or a copy-out
        {
          std::lock_guard<std::mutex> lk(mu_);
          Job& j = jobs_[tk];
          j.state = 1;
          j.t.start_s = now();
        }
        p->start(r);
        sl.ticket = tk;
        sl.active = p;
        sl.active_key = k;
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
        if (f.err.empty() && f.len > j->cap) f.err = "server: proof larger than the output buffer";
        if (f.err.empty()) std::memcpy(j->out, f.src, f.len);
        std::lock_guard<std::mutex> lk(mu_);
        j->len = f.err.empty() ? f.len : 0;
        j->t.done_s = f.t;
        j->err = f.err;
        j->state = f.err.empty() ? 2 : 3;
      }
      cv_done_.notify_all();
    }
    for (auto& sl : slots_) sl.pending = false;  // every finished proof has been copied
    if (!progress) std::this_thread::yield();
  }
}

}  // namespace vx
