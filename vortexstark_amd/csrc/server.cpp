// HIP provers for the proof server's scheduler (see server.h, server_core.h).
#include "server.h"

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

// A resident prover of one size class on a slot's stream. Its completion
// event is recorded behind each proof (graph replay + download kernel).
class Served final : public ServedProver {
 public:
  Served(const ProofRequest& r, hipStream_t s, bool graphs)
      : p_(r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, s, air_of(r)) {
    VX_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    p_.set_graphs(graphs);
    if (graphs) p_.prepare_graph();
  }
  ~Served() override {
    (void)hipStreamSynchronize(p_.stream());
    if (done_) (void)hipEventDestroy(done_);
  }
  void start(const ProofRequest& r) override {
    if (r.air == AIR_D2) p_.enqueue((uint32_t)r.seed, (uint32_t)(r.seed >> 32), false);
    else p_.enqueue(r.a, r.b, false);
    VX_HIP(hipEventRecord(done_, p_.stream()));
  }
  int poll(std::string* err) override {
    const hipError_t e = hipEventQuery(done_);
    if (e == hipErrorNotReady) return 0;
    if (e == hipSuccess) return 1;
    if (err) *err = hipGetErrorString(e);
    return -1;
  }
  const uint8_t* result() const override { return p_.host_proof(); }
  size_t result_bytes() const override { return p_.proof_bytes(); }
  uint64_t device_bytes() const override { return p_.device_bytes(); }

 private:
  StarkProver p_;
  hipEvent_t done_ = nullptr;
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

ProofServer::ProofServer(uint32_t n_slots, uint64_t mem_budget_bytes, bool graphs) : graphs_(graphs) {
  require(n_slots >= 1 && n_slots <= 64, "server: n_slots must be 1..64");
  VX_HIP(hipGetDevice(&device_));
  streams_.assign(n_slots, nullptr);
  try {
    for (auto& s : streams_) VX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ServerHooks h;
    h.create = [this](uint32_t slot, const ProofRequest& r) -> std::unique_ptr<ServedProver> {
      VX_HIP(hipSetDevice(device_));
      return std::make_unique<Served>(r, streams_[slot], graphs_);
    };
    h.prover_bytes = [](const ProofRequest& r) {
      return StarkProver::estimate_bytes(r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, air_of(r));
    };
    h.table_bytes = [](const ProofRequest& r) { return ProverTables::estimate_bytes(r.log_n, r.blowup_bits); };
    h.thread_init = [this] { VX_HIP(hipSetDevice(device_)); };
    core_ = std::make_unique<ServerCore>(n_slots, mem_budget_bytes, std::move(h));
  } catch (...) {
    for (auto& s : streams_)
      if (s) (void)hipStreamDestroy(s);
    throw;
  }
}

ProofServer::~ProofServer() {
  core_.reset();  // drains the queue and destroys the provers, which use the streams
  for (auto& s : streams_)
    if (s) (void)hipStreamDestroy(s);
}

void ProofServer::prepare(const ProofRequest& r) {
  check_request(r);
  VX_HIP(hipSetDevice(device_));
  core_->prepare(r);
}

uint64_t ProofServer::submit(const ProofRequest& r, uint8_t* out, size_t cap) {
  check_request(r);
  require(out != nullptr, "server: null output buffer");
  const size_t need = proof_size(r.log_n, r.blowup_bits, r.n_queries, r.pow_bits, air_of(r));
  require(cap >= need, "server: output buffer smaller than the proof (see vx_proof_size)");
  return core_->submit(r, out, cap);
}

}  // namespace vx
