// Autotuner for the Merkle subtree schedule.
//
// A commitment's speed depends on its tile shape (leaves per workgroup and
// per lane, TreeShape) in ways that change with the leaf kind, the number of
// columns and the tree size: small trees are latency-bound and want one leaf
// per lane, large single-column trees want 16 leaves per lane, multi-column
// leaves and FRI folds want 4 (coalesced loads). Instead of trusting one
// table, the first commitment of each key (leaf kind, columns, constraints,
// log_n, stored levels) on a device is timed with every valid shape, twice
// each (the second run counts), on the caller's stream and with the
// caller's buffers, and the fastest is cached for the process (the built-in
// shape unless another is more than 3 % faster). All shapes
// produce the same tree, so the choice never changes a proof. Measuring
// needs a stream synchronisation, so it never happens while the stream is
// being captured into a graph (the provers run one eager proof first, see
// StarkProver::tune). An optional cache file keeps the choices across
// processes: one line "kind n_cols n_cons lift_n log_n store_lo tile_log lpt"
// per key.
//
// The circle NTT is tuned the same way per transform shape (direction,
// log_n, transformed bits, columns, basis): the built-in plan, tiles in
// launch order instead of XCD-contiguous ranges, and (where the strided bits
// allow) the strided bits over one more pass, each timed twice on the
// caller's stream; out-of-place transforms are timed on the caller's
// buffers, in-place ones on a scratch buffer of the same shape. Cache-file
// lines "ntt fwd log_n eff_bits n_cols stwo split order".
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "tune.h"
#include "vx_internal.h"

namespace vx {

namespace {

using Key = std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;

using NKey = std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;

struct Tuner {
  std::mutex mu;
  bool enabled = true;
  std::string file;
  std::map<Key, TreeShape> cache;
  std::map<NKey, NttPlan> ntt;
  uint64_t measured = 0;
};

NKey nkey_of(int dev, const NttKey& k) {
  return NKey{dev, k.fwd ? 1u : 0u, k.log_n, k.eff_bits, k.n_cols, k.stwo ? 1u : 0u};
}

Tuner& tuner() {
  static Tuner t;
  return t;
}

Key key_of(int dev, const LeafSrc& L, uint32_t log_n, uint32_t store_lo) {
  const bool cols = L.kind == LEAF_COLS || L.kind == LEAF_D2_QUOTIENT;
  return Key{dev,
             L.kind,
             cols ? L.n_cols : 0u,
             L.kind == LEAF_D2_QUOTIENT ? L.n_cons : 0u,
             (L.kind == LEAF_LIFTED || L.kind == LEAF_LIFTED_RLC) ? L.lift_n : 0u,
             log_n,
             store_lo};
}

void append_line(const std::string& file, const Key& k, const TreeShape& s) {
  if (file.empty()) return;
  FILE* f = std::fopen(file.c_str(), "a");
  if (!f) return;  // the cache file is an optimisation: a read-only path only costs re-measuring
  std::fprintf(f, "%u %u %u %u %u %u %u %u\n", (unsigned)std::get<1>(k), std::get<2>(k), std::get<3>(k),
               std::get<4>(k), std::get<5>(k), std::get<6>(k), s.tile_log, s.lpt);
  std::fclose(f);
}

// A candidate replaces the built-in shape only when it is faster by more
// than this fraction, so that near-ties (run-to-run noise) keep the same,
// reproducible choice in every process.
constexpr float MIN_GAIN = 0.03f;

TreeShape measure(const LeafSrc& L, const TreeLevels& T, uint32_t store_lo, hipStream_t s,
                  const std::vector<TreeShape>& cands, const TreeShape& dflt) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VX_HIP(hipEventCreate(&e0));
  VX_HIP(hipEventCreate(&e1));
  TreeShape best = dflt;
  float best_ms = 1e30f, dflt_ms = 1e30f;
  try {
    for (const TreeShape& c : cands) {
      float ms = 0;
      for (int rep = 0; rep < 2; rep++) {
        VX_HIP(hipEventRecord(e0, s));
        merkle_commit_shaped(L, T, store_lo, RootHook{}, s, c);
        VX_HIP(hipEventRecord(e1, s));
        VX_HIP(hipEventSynchronize(e1));
        VX_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      if (c.tile_log == dflt.tile_log && c.lpt == dflt.lpt) dflt_ms = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    if (dflt_ms <= best_ms * (1.0f + MIN_GAIN)) best = dflt;
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  VX_HIP(hipEventDestroy(e0));
  VX_HIP(hipEventDestroy(e1));
  return best;
}

}  // namespace

TreeShape tuned_shape(const LeafSrc& L, const TreeLevels& T, uint32_t store_lo, hipStream_t s) {
  TreeShape dflt = merkle_default_shape(L, T.log_n);
  const std::vector<TreeShape> cands = merkle_shapes(T.log_n, store_lo);
  // a caller storing lower levels than the built-in shape allows gets the
  // first shape that does (merkle_commit rejects the call if none does)
  if (!cands.empty() && !(dflt.lpt <= store_lo && store_lo <= dflt.tile_log)) dflt = cands[0];
  Tuner& t = tuner();
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  const Key k = key_of(dev, L, T.log_n, store_lo);
  {
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.enabled) return dflt;
    auto it = t.cache.find(k);
    if (it != t.cache.end()) return it->second;
  }
  if (cands.size() < 2) return dflt;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  VX_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return dflt;
  const TreeShape best = measure(L, T, store_lo, s, cands, dflt);
  std::lock_guard<std::mutex> g(t.mu);
  auto ins = t.cache.emplace(k, best);
  if (ins.second) {
    t.measured++;
    append_line(t.file, k, best);
  }
  return ins.first->second;
}

NttPlan known_ntt_plan(const NttKey& key) {
  Tuner& t = tuner();
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(t.mu);
  if (!t.enabled) return NttPlan{};
  auto it = t.ntt.find(nkey_of(dev, key));
  return it == t.ntt.end() ? NttPlan{} : it->second;
}

NttPlan tuned_ntt_plan(const NttKey& key, hipStream_t s, const std::function<bool(const NttPlan&)>& run) {
  Tuner& t = tuner();
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  const NKey k = nkey_of(dev, key);
  {
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.enabled) return NttPlan{};
    auto it = t.ntt.find(k);
    if (it != t.ntt.end()) return it->second;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  VX_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return NttPlan{};
  const uint32_t contig = key.eff_bits < 14 ? key.eff_bits : 14;
  const uint32_t strided = key.eff_bits - contig;
  std::vector<NttPlan> cands = {NttPlan{0, 0}, NttPlan{0, 1}};
  if (strided >= 2) cands.push_back(NttPlan{1, 0});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VX_HIP(hipEventCreate(&e0));
  VX_HIP(hipEventCreate(&e1));
  NttPlan best{};
  try {
    float best_ms = 1e30f, dflt_ms = 1e30f;
    bool timed = true;
    for (const NttPlan& c : cands) {
      float ms = 0;
      for (int rep = 0; rep < 2 && timed; rep++) {
        VX_HIP(hipEventRecord(e0, s));
        timed = run(c);
        VX_HIP(hipEventRecord(e1, s));
        VX_HIP(hipEventSynchronize(e1));
        VX_HIP(hipEventElapsedTime(&ms, e0, e1));
      }
      if (!timed) break;
      if (c.split == 0 && c.order == 0) dflt_ms = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    if (!timed || dflt_ms <= best_ms * (1.0f + MIN_GAIN)) best = NttPlan{};
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  VX_HIP(hipEventDestroy(e0));
  VX_HIP(hipEventDestroy(e1));
  std::lock_guard<std::mutex> g(t.mu);
  auto ins = t.ntt.emplace(k, best);
  if (ins.second) {
    t.measured++;
    if (!t.file.empty()) {
      if (FILE* f = std::fopen(t.file.c_str(), "a")) {
        std::fprintf(f, "ntt %u %u %u %u %u %u %u\n", std::get<1>(k), std::get<2>(k), std::get<3>(k),
                     std::get<4>(k), std::get<5>(k), best.split, best.order);
        std::fclose(f);
      }
    }
  }
  return ins.first->second;
}

std::vector<uint32_t> tuner_ntt_entries() {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  std::vector<uint32_t> out;
  for (const auto& kv : t.ntt) {
    const NKey& k = kv.first;
    const uint32_t e[8] = {(uint32_t)std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k),
                           std::get<4>(k), std::get<5>(k), kv.second.split, kv.second.order};
    out.insert(out.end(), e, e + 8);
  }
  return out;
}

void tuner_configure(bool enabled, const char* cache_file) {
  Tuner& t = tuner();
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(t.mu);
  t.enabled = enabled;
  t.file = cache_file ? cache_file : "";
  if (t.file.empty()) return;
  FILE* f = std::fopen(t.file.c_str(), "r");
  if (!f) return;
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    unsigned v[8];
    if (std::sscanf(line, "ntt %u %u %u %u %u %u %u", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6]) == 7) {
      if (v[5] <= 1 && v[6] <= 1) t.ntt[NKey{dev, v[0], v[1], v[2], v[3], v[4]}] = NttPlan{v[5], v[6]};
      continue;
    }
    if (std::sscanf(line, "%u %u %u %u %u %u %u %u", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]) != 8)
      continue;
    const TreeShape sh{v[6], v[7]};
    // keep only shapes this build accepts for that key
    bool ok = false;
    for (const TreeShape& c : merkle_shapes(v[4], v[5])) ok = ok || (c.tile_log == sh.tile_log && c.lpt == sh.lpt);
    if (ok) t.cache[Key{dev, v[0], v[1], v[2], v[3], v[4], v[5]}] = sh;
  }
  std::fclose(f);
}

std::vector<TunedEntry> tuner_entries() {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  std::vector<TunedEntry> out;
  for (const auto& kv : t.cache) {
    const Key& k = kv.first;
    out.push_back({(uint32_t)std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k), std::get<4>(k),
                   std::get<5>(k), std::get<6>(k), kv.second.tile_log, kv.second.lpt});
  }
  return out;
}

}  // namespace vx
