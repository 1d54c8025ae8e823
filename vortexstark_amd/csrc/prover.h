// Host-side prover over the gfx950 kernels: the C++ counterpart of
// src/prover.rs (prove / prove_cached / ProverCache). One StarkProver owns
// the twiddles, every intermediate buffer and the proof buffer for one size
// class (AIR, log_n, blowup, queries, pow), so repeated proofs allocate
// nothing and run as a single stream of kernels ending in one D2H copy.
// AIRs: the reference's Fibonacci AIR (KRAK v1/v2 bytes) and the synthetic
// multi-column degree-2 AIR of the serving workload (vx_d2.h, KRAK v3).
#pragma once
#include <memory>
#include <vector>

#include "vx_internal.h"

namespace vx {

struct DeviceBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DeviceBuf() = default;
  explicit DeviceBuf(size_t n);
  ~DeviceBuf();
  DeviceBuf(const DeviceBuf&) = delete;
  DeviceBuf& operator=(const DeviceBuf&) = delete;
  DeviceBuf(DeviceBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DeviceBuf& operator=(DeviceBuf&& o) noexcept;
  uint32_t* u32() const { return static_cast<uint32_t*>(p); }
};

// AIR_BATCH: the batched-columns proof (KRAK v4): columns of uneven log
// sizes (sorted ascending, 14..24), bit-reversed LDE, one lifted tree, a
// random-linear-combination quotient, then the Fibonacci FRI.
enum AirKind : uint32_t { AIR_FIBONACCI = 0, AIR_D2 = 1, AIR_BATCH = 2 };
struct AirSpec {
  uint32_t kind = AIR_FIBONACCI;
  uint32_t n_cols = 1;    // AIR_D2: trace columns W; AIR_BATCH: columns
  uint32_t n_cons = 1;    // AIR_D2: constraints K
  uint64_t air_seed = 0;  // AIR_D2: constraint coefficient seed
  std::vector<uint32_t> log_sizes;  // AIR_BATCH: per-column log sizes
};

// Word offsets of every variable field of a proof: KRAK v1/v2
// (src/bin/stark_cli.rs:1092-1120) for Fibonacci, KRAK v3 for AIR_D2
// (header, one root and one next-row decommitment per trace group).
struct ProofLayout {
  uint32_t log_n, blowup_bits, nq, pow_bits, le, n_fri, n_groups;
  uint64_t header_words, w_qroot, w_nfri, w_froot0, w_last, w_nq, w_qdec, w_nfdec, w_nonce;
  std::vector<uint64_t> w_troot, w_tdec, w_fdec;
  uint64_t total_words;
  ProofLayout(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits, const AirSpec& air = {});
};

// Reference-layout twiddle tables of one coset, resident on the device.
struct TwiddleSet {
  DeviceBuf line, circle;
  TwiddleRef ref{};
  // forward tables of `c`; `inverse` replaces them with their inverses
  TwiddleSet(const coset& c, bool inverse, hipStream_t s);
};

struct StageTimes {
  float trace_ms, interpolate_ms, lde_ms, trace_commit_ms, quotient_commit_ms, fri_ms, queries_ms, decommit_ms,
      download_ms, total_ms, trace_subtree_ms;
};

// Read-only tables of one (log_n, blowup) size class: interpolation twiddles
// (inverse, trace domain), evaluation twiddles (eval domain) and the inverse
// FRI fold twiddles of every layer. Shared by every prover of the class on a
// device (ProverTables::get), built once.
struct ProverTables {
  uint32_t log_n, le, n_fri;
  std::unique_ptr<TwiddleSet> trace_itw, eval_tw;
  DeviceBuf fold_tw;
  std::vector<uint64_t> fold_tw_off;
  uint64_t bytes = 0;
  ProverTables(uint32_t log_n, uint32_t blowup_bits, hipStream_t s);
  static std::shared_ptr<const ProverTables> get(uint32_t log_n, uint32_t blowup_bits, hipStream_t s);
  static uint64_t estimate_bytes(uint32_t log_n, uint32_t blowup_bits);
};

class StarkProver {
 public:
  StarkProver(uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits, hipStream_t stream,
              const AirSpec& air = {});
  ~StarkProver();
  // Enqueue one proof on the stream (no host synchronisation): Fib(a, b),
  // or for AIR_D2 the trace of seed (a | b << 32). Untimed proofs replay a
  // HIP graph captured on the first one (when graphs are enabled); the
  // public inputs reach it through device memory.
  void enqueue(uint32_t a, uint32_t b, bool timed = false);
  // Wait for the last enqueued proof; returns its KRAK bytes.
  const uint8_t* wait(size_t* len);
  size_t proof_bytes() const { return layout_.total_words * 4; }
  const ProofLayout& layout() const { return layout_; }
  hipStream_t stream() const { return stream_; }
  uint64_t device_bytes() const { return device_bytes_; }
  StageTimes last_times();
  void set_graphs(bool on) { use_graph_ = on; }
  // Capture the proof graph now rather than on the first untimed proof
  // (after one eager proof that lets the autotuner time every commitment).
  void prepare_graph() {
    if (use_graph_ && !graph_) {
      tune();
      capture();
    }
  }
  // One eager proof of dummy inputs: every merkle_commit of this class is
  // timed by the autotuner (if enabled) before a graph freezes the shapes.
  void tune();
  bool graphs() const { return use_graph_; }
  // Device bytes a prover of this class allocates for itself (tables excluded).
  static uint64_t estimate_bytes(uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits,
                                 const AirSpec& air = {});
  const AirSpec& air() const { return air_; }

  // device buffers for tests / introspection
  const uint32_t* d_eval() const { return eval_.u32(); }
  const uint32_t* d_trace() const { return trace_.u32(); }
  const uint32_t* d_proof() const { return proof_.u32(); }
  // host (pinned) copy of the last proof; two buffers alternate, so the bytes
  // of proof k stay valid while proof k + 1 runs
  const uint8_t* host_proof() const { return host_proof_[cur_]; }
  const uint32_t* d_layer(uint32_t k) const { return layers_[k].u32(); }

 private:
  void build_descs();
  void build_batch();
  void record(bool timed, hipStream_t s);
  void capture();
  TreeLevels tree_levels(uint32_t idx) const;

  AirSpec air_;
  ProofLayout layout_;
  uint32_t log_n_, blowup_, nq_, pow_bits_, le_, n_fri_, n_groups_;
  hipStream_t stream_;
  bool own_stream_ = false;
  uint64_t device_bytes_ = 0;

  std::shared_ptr<const ProverTables> tab_;
  DeviceBuf trace_, eval_;         // W columns of 2^log_n / 2^le (W = 1 for Fibonacci)
  DeviceBuf quot_, cons_, beta_;   // AIR_D2: stored quotient (SoA), constraint table, challenge powers
  // AIR_BATCH: coefficients, bit-reversed LDE, work lists, lifted columns,
  // column offsets (trace_ holds the inputs, eval_ the natural-order LDE)
  DeviceBuf coef_, evr_, lde_jobs_, br_jobs_, lift_, fill_off_, size_digest_;
  std::vector<std::shared_ptr<const ProverTables>> col_tables_;
  LdeBatchStages lde_st_{};
  uint32_t br_count_ = 0;
  uint64_t br_tiles_ = 0, n_in_ = 0;
  std::vector<DeviceBuf> layers_;  // FRI layer k: 4 SoA columns of 2^(le-1-k)
  std::vector<uint64_t> layer_col_words_;
  std::vector<DeviceBuf> trees_;   // g < G trace groups, G = quotient, G + 1 + k = FRI layer k
  std::vector<uint32_t> tree_log_, tree_lo_;
  std::vector<std::vector<uint64_t>> tree_off_;
  DeviceBuf small_;                // channel | alphas | queries | prefix | nonce | inputs
  Channel* d_channel_ = nullptr;
  uint32_t* d_alphas_ = nullptr;
  uint32_t* d_queries_ = nullptr;
  uint32_t* d_prefix_ = nullptr;
  uint64_t* d_nonce_ = nullptr;
  uint32_t* d_inputs_ = nullptr;
  DeviceBuf descs_;
  DeviceBuf proof_;
  uint8_t* host_proof_[2] = {nullptr, nullptr};
  uint32_t* host_proof_dev_[2] = {nullptr, nullptr};  // the same pinned buffers as the device addresses them
  uint32_t cur_ = 0;
  hipEvent_t ev_[12] = {};
  bool timed_ = false;
  bool use_graph_ = true;
  bool tuned_ = false;
  hipGraphExec_t graph_ = nullptr;
};

// Host verifier (src/verifier.rs:39-310) over KRAK bytes; returns true and
// leaves `err` empty when the proof is accepted.
bool verify_proof(const uint8_t* bytes, size_t len, uint32_t blowup_bits, uint32_t n_queries, std::string& err);
size_t proof_size(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits, const AirSpec& air = {});

}  // namespace vx
