// Host-only tests of the proof server's scheduler (server_core.h) with mock
// provers whose completion the test controls. Built with AddressSanitizer by
// tests/test_server_core.py: a proof copied out of a prover that was already
// destroyed is reported as a use-after-free, and the mock also poisons its
// buffers on destruction, so the bytes would not match either.
//
// Usage: server_core_test [scenario]; prints "PASS <name>" per scenario and
// exits non-zero on the first failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "server_core.h"
#include "vx_error.h"

using namespace vx;

#define CHECK(c)                                                                   \
  do {                                                                             \
    if (!(c)) {                                                                    \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);          \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

namespace {

struct Control {
  std::atomic<bool> hold{false};        // while set, running proofs do not finish
  std::atomic<int> started{0};          // proofs started
  std::atomic<int> alive{0};            // mock provers alive
  std::atomic<int> fail_poll{0};        // the next n completions report a device error
  std::atomic<int> fail_start{0};       // the next n starts throw
  std::atomic<int> running{0};          // proofs in flight
  std::atomic<int> max_running{0};
};
Control ctl;

size_t proof_len(const ProofRequest& r) { return 64 + 8 * r.log_n + r.pow_bits; }
uint8_t proof_byte(const ProofRequest& r, size_t i) {
  return (uint8_t)(r.a * 131u + r.b * 7u + r.log_n * 17u + r.pow_bits * 3u + i * 29u + (i >> 8));
}
uint64_t prover_bytes(const ProofRequest& r) { return 1000ull << (r.log_n - 10); }
uint64_t table_bytes(const ProofRequest& r) { return 100ull << (r.log_n - 10); }

class MockProver final : public ServedProver {
 public:
  explicit MockProver(const ProofRequest& r) : len_(proof_len(r)), bytes_(prover_bytes(r)) {
    for (auto& b : buf_) b = new uint8_t[len_];
    ctl.alive++;
  }
  ~MockProver() override {
    for (auto& b : buf_) {
      std::memset(b, 0xDD, len_);
      delete[] b;
    }
    ctl.alive--;
  }
  void start(const ProofRequest& r) override {
    CHECK(!running_);
    if (ctl.fail_start > 0) {
      ctl.fail_start--;
      throw Error(ERR_HIP, "mock: launch failed");
    }
    req_ = r;
    running_ = true;
    cur_ ^= 1;
    ctl.started++;
    int now = ++ctl.running;
    int prev = ctl.max_running.load();
    while (now > prev && !ctl.max_running.compare_exchange_weak(prev, now)) {
    }
  }
  int poll(std::string* err) override {
    CHECK(running_);
    if (ctl.hold) return 0;
    running_ = false;
    ctl.running--;
    if (ctl.fail_poll > 0) {
      ctl.fail_poll--;
      *err = "mock: memory access fault";
      return -1;
    }
    for (size_t i = 0; i < len_; i++) buf_[cur_][i] = proof_byte(req_, i);  // the proof "lands" now
    return 1;
  }
  const uint8_t* result() const override { return buf_[cur_]; }
  size_t result_bytes() const override { return len_; }
  uint64_t device_bytes() const override { return bytes_; }

 private:
  size_t len_;
  uint64_t bytes_;
  uint8_t* buf_[2];
  int cur_ = 0;
  bool running_ = false;
  ProofRequest req_{};
};

ServerHooks hooks() {
  ServerHooks h;
  h.create = [](uint32_t, const ProofRequest& r) -> std::unique_ptr<ServedProver> {
    return std::make_unique<MockProver>(r);
  };
  h.prover_bytes = prover_bytes;
  h.table_bytes = table_bytes;
  return h;
}

ProofRequest req(uint32_t log_n, uint32_t a, uint32_t pow = 0) {
  ProofRequest r{};
  r.air = 0;
  r.log_n = log_n;
  r.blowup_bits = 2;
  r.n_queries = 80;
  r.pow_bits = pow;
  r.a = a;
  r.b = a + 1;
  return r;
}

struct Out {
  ProofRequest r;
  std::vector<uint8_t> buf;
  uint64_t ticket = 0;
};

uint64_t submit(ServerCore& s, Out& o) {
  o.buf.assign(proof_len(o.r) + 16, 0);
  o.ticket = s.submit(o.r, o.buf.data(), o.buf.size());
  return o.ticket;
}

void expect_ok(ServerCore& s, Out& o) {
  size_t len = 0;
  JobTimes t{};
  std::string err;
  bool ok = s.wait(o.ticket, &len, &t, err);
  if (!ok) std::fprintf(stderr, "unexpected failure: %s\n", err.c_str());
  CHECK(ok);
  CHECK(len == proof_len(o.r));
  for (size_t i = 0; i < len; i++) CHECK(o.buf[i] == proof_byte(o.r, i));
  CHECK(t.submit_s <= t.start_s && t.start_s <= t.done_s);
}

void wait_until(const std::function<bool()>& f) {
  for (int i = 0; i < 20000 && !f(); i++) std::this_thread::sleep_for(std::chrono::microseconds(100));
  CHECK(f());
}

void settle() { std::this_thread::sleep_for(std::chrono::milliseconds(30)); }

// The interleaving behind a use-after-free: a slot's proof finishes while a
// request of a new class waits, and the only eviction candidate is the
// prover that holds the finished, not yet copied proof.
void evict_while_pending() {
  ctl.hold = true;
  const int s0 = ctl.started;
  {
    ServerCore s(1, prover_bytes(req(10, 0)) + table_bytes(req(10, 0)), hooks());
    Out x{req(10, 5)}, y{req(10, 9, 7)};  // same shape, different class (PoW)
    submit(s, x);
    wait_until([&] { return ctl.started == s0 + 1; });
    submit(s, y);
    settle();  // y is in the dispatcher's queue, waiting for the slot
    ctl.hold = false;
    expect_ok(s, x);
    expect_ok(s, y);
    ServerStats st = s.stats();
    CHECK(st.deferred >= 1);
    CHECK(st.evicted == 1);
    CHECK(st.created == 2);
    CHECK(s.resident_bytes() == s.recount_bytes());
    CHECK(s.resident_bytes() <= prover_bytes(req(10, 0)) + table_bytes(req(10, 0)));
  }
  CHECK(ctl.alive == 0);
}

// Two slots, a budget of about two provers, classes alternating, random
// completion order: every proof equals its request's bytes.
void two_slots_alternating() {
  const uint64_t budget = 2 * prover_bytes(req(12, 0)) + 2 * table_bytes(req(12, 0));
  std::mt19937 rng(3);
  {
    ServerCore s(2, budget, hooks());
    std::vector<Out> outs;
    const uint32_t sizes[] = {12, 10, 11, 12, 10, 11};
    for (int i = 0; i < 60; i++) outs.push_back(Out{req(sizes[i % 6], (uint32_t)i, (i / 6) % 2 ? 4 : 0)});
    for (size_t i = 0; i < outs.size(); i++) {
      ctl.hold = (rng() % 3) == 0;
      submit(s, outs[i]);
      if (rng() % 4 == 0) settle();
      ctl.hold = false;
    }
    for (auto& o : outs) expect_ok(s, o);
    CHECK(s.resident_bytes() <= budget);
    CHECK(s.resident_bytes() == s.recount_bytes());
    CHECK(s.stats().evicted > 0);
    CHECK(ctl.max_running <= 2);
  }
  CHECK(ctl.alive == 0);
}

// More shapes than the budget holds, cycled: the tables of an evicted shape
// are released, so nothing is refused.
void cycle_more_classes_than_fit() {
  const uint64_t budget = prover_bytes(req(14, 0)) + table_bytes(req(14, 0));
  {
    ServerCore s(1, budget, hooks());
    for (int round = 0; round < 3; round++)
      for (uint32_t ln = 10; ln <= 14; ln++) {
        Out o{req(ln, ln * 10 + round)};
        submit(s, o);
        expect_ok(s, o);
        CHECK(s.resident_bytes() == s.recount_bytes());
        CHECK(s.resident_bytes() <= budget);
      }
    // the 2^14 class fills the budget alone: every visit to it evicted all
    // other shapes, their tables included
    CHECK(s.stats().evicted >= 6);
  }
  CHECK(ctl.alive == 0);
}

// Classes of one (log_n, blowup) share their tables: counted once.
void shared_tables_counted_once() {
  const ProofRequest a = req(12, 1, 0), b = req(12, 2, 10);
  const uint64_t budget = 4 * prover_bytes(a) + table_bytes(a);
  {
    ServerCore s(2, budget, hooks());
    s.prepare(a);
    s.prepare(b);
    CHECK(s.resident_bytes() == budget);
    CHECK(s.resident_bytes() == s.recount_bytes());
    Out o{b};
    submit(s, o);
    expect_ok(s, o);
    CHECK(s.stats().created == 4);
    bool threw = false;
    try {
      s.prepare(req(11, 0));
    } catch (const Error& e) {
      threw = std::strstr(e.what(), "budget") != nullptr;
    }
    CHECK(threw);
  }
  CHECK(ctl.alive == 0);
}

void rejects_what_cannot_fit() {
  {
    ServerCore s(2, 1500, hooks());
    Out big{req(12, 1)}, small{req(10, 2)};
    submit(s, big);
    submit(s, small);
    size_t len;
    JobTimes t;
    std::string err;
    CHECK(!s.wait(big.ticket, &len, &t, err));
    CHECK(err.find("budget") != std::string::npos);
    expect_ok(s, small);
  }
  CHECK(ctl.alive == 0);
}

void device_and_launch_failures() {
  {
    ServerCore s(2, 1 << 20, hooks());
    Out a{req(10, 1)}, b{req(10, 2)}, c{req(11, 3)}, d{req(11, 4)};
    ctl.fail_poll = 1;
    submit(s, a);
    size_t len;
    JobTimes t;
    std::string err;
    CHECK(!s.wait(a.ticket, &len, &t, err));
    CHECK(err.find("fault") != std::string::npos);
    submit(s, b);
    expect_ok(s, b);
    ctl.fail_start = 1;
    submit(s, c);
    CHECK(!s.wait(c.ticket, &len, &t, err));
    CHECK(err.find("launch") != std::string::npos);
    submit(s, d);
    expect_ok(s, d);
  }
  CHECK(ctl.alive == 0);
}

// prepare() on another thread while requests of other classes flow through.
void prepare_while_serving() {
  {
    ServerCore s(3, 1ull << 30, hooks());
    std::vector<Out> outs;
    for (int i = 0; i < 40; i++) outs.push_back(Out{req(10 + i % 3, (uint32_t)i)});
    std::thread prep([&] {
      for (uint32_t ln = 13; ln <= 16; ln++) s.prepare(req(ln, 0));
    });
    for (auto& o : outs) submit(s, o);
    prep.join();
    for (auto& o : outs) expect_ok(s, o);
    CHECK(s.resident_bytes() == s.recount_bytes());
  }
  CHECK(ctl.alive == 0);
}

// The destructor finishes what is queued before it stops.
void destructor_drains() {
  std::vector<Out> outs;
  for (int i = 0; i < 12; i++) outs.push_back(Out{req(10 + i % 2, (uint32_t)i)});
  {
    ServerCore s(2, 1 << 20, hooks());
    for (auto& o : outs) submit(s, o);
  }
  for (auto& o : outs)
    for (size_t i = 0; i < proof_len(o.r); i++) CHECK(o.buf[i] == proof_byte(o.r, i));
  CHECK(ctl.alive == 0);
}

}  // namespace

int main(int argc, char** argv) {
  const std::map<std::string, void (*)()> all = {
      {"evict_while_pending", evict_while_pending},
      {"two_slots_alternating", two_slots_alternating},
      {"cycle_more_classes_than_fit", cycle_more_classes_than_fit},
      {"shared_tables_counted_once", shared_tables_counted_once},
      {"rejects_what_cannot_fit", rejects_what_cannot_fit},
      {"device_and_launch_failures", device_and_launch_failures},
      {"prepare_while_serving", prepare_while_serving},
      {"destructor_drains", destructor_drains},
  };
  for (auto& kv : all) {
    if (argc > 1 && kv.first != argv[1]) continue;
    kv.second();
    std::printf("PASS %s\n", kv.first.c_str());
    std::fflush(stdout);
  }
  return 0;
}
