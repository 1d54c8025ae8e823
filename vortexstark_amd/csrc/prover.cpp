// Fibonacci Circle-STARK prover driver (src/prover.rs:172-619 counterpart).
//
// The transcript order is the reference's (prover.rs:321-557):
//   commit trace LDE -> mix root, draw alpha -> commit quotient -> mix root,
//   draw alpha_0 -> circle fold, commit layer 0 -> mix -> { draw alpha_k,
//   line fold, commit, mix } down to 8 values -> mix_felts(last layer) ->
//   [PoW: grind, mix_u64(nonce)] -> 80 query draws -> decommit.
// Everything, transcript included, runs on one HIP stream; the only host
// transfer is the finished KRAK proof.
#include "prover.h"

#include "vx_d2.h"

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

namespace vx {

static constexpr uint32_t KRAK_MAGIC = 0x4B52414Bu;

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? ERR_ALLOC : ERR_HIP;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}

DeviceBuf::DeviceBuf(size_t n) : bytes(n) {
  if (n) VX_HIP(hipMalloc(&p, n));
}
DeviceBuf::~DeviceBuf() {
  if (p) (void)hipFree(p);
}
DeviceBuf& DeviceBuf::operator=(DeviceBuf&& o) noexcept {
  if (this != &o) {
    if (p) (void)hipFree(p);
    p = o.p;
    bytes = o.bytes;
    o.p = nullptr;
    o.bytes = 0;
  }
  return *this;
}

ProofLayout::ProofLayout(uint32_t log_n_, uint32_t blowup_bits_, uint32_t nq_, uint32_t pow_bits_,
                         const AirSpec& air)
    : log_n(log_n_), blowup_bits(blowup_bits_), nq(nq_), pow_bits(pow_bits_) {
  le = log_n + blowup_bits;
  n_fri = le - 3;
  const bool d2 = air.kind == AIR_D2, batch = air.kind == AIR_BATCH;
  n_groups = d2 ? d2_groups(air.n_cols) : 1;
  // v1: magic, version, log_n, a, b; v2 adds blowup, pow, nonce lo/hi;
  // v3: magic, 3, log_n, W, K, blowup, pow, nonce lo/hi, seed lo/hi,
  //     air_seed lo/hi, G;  v4: magic, 4, C, blowup, pow, nonce lo/hi,
  //     seed lo/hi, C log sizes
  header_words = batch ? 9 + air.n_cols : (d2 ? 14 : (pow_bits ? 9 : 5));
  w_nonce = batch ? 5 : (d2 ? 7 : (pow_bits ? 7 : 0));
  uint64_t w = header_words;
  for (uint32_t g = 0; g < n_groups; g++) {
    w_troot.push_back(w);
    w += 8;
  }
  w_qroot = w;
  w_nfri = w_qroot + 8;
  w_froot0 = w_nfri + 1;
  w_last = w_froot0 + 8ull * n_fri;
  w_nq = w_last + 1 + 32;
  w = w_nq + 1 + nq;
  for (uint32_t g = 0; g < n_groups; g++) {
    const uint64_t width = d2 ? d2_group_cols(air.n_cols, g) : (batch ? air.n_cols : 1);
    w_tdec.push_back(w);
    w += 1 + 2ull * nq * width + 2ull * nq * (1 + 8ull * le);
  }
  w_qdec = w;
  w_nfdec = w_qdec + 1 + 8ull * nq + 2ull * nq * (1 + 8ull * le);
  w = w_nfdec + 1;
  for (uint32_t k = 0; k < n_fri; k++) {
    w_fdec.push_back(w);
    uint32_t L = le - 1 - k;
    w += 1 + 8ull * nq + 2ull * nq * (1 + 8ull * L);
  }
  total_words = w;
}

size_t proof_size(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits, const AirSpec& air) {
  return ProofLayout(log_n, blowup_bits, nq, pow_bits, air).total_words * 4;
}

// src/circle.rs:284-347: coset points, circle twiddles = y of the first half,
// line layer l = even x of layer l-1's squashed x (2x^2 - 1).
TwiddleSet::TwiddleSet(const coset& c, bool inverse, hipStream_t s) {
  const uint32_t L = c.log_size;
  const uint64_t n = 1ull << L;
  line = DeviceBuf(4 * (n > 2 ? n : 2));
  circle = DeviceBuf(4 * (n / 2 > 0 ? n / 2 : 1));
  DeviceBuf xs(4 * n), ys(4 * n), tmp(4 * (n / 2 > 0 ? n / 2 : 1));
  coset_points(c.init, c.step, xs.u32(), ys.u32(), n, s);
  VX_HIP(hipMemcpyAsync(circle.p, ys.p, 4 * (n / 2), hipMemcpyDeviceToDevice, s));
  uint64_t off = 0, cur_n = n;
  uint32_t* cur = xs.u32();
  uint32_t* nxt = tmp.u32();
  for (uint32_t l = 0; l + 1 < L; l++) {
    uint64_t half = cur_n / 2;
    ref.off[l] = (uint32_t)off;
    extract_and_squash(cur, line.u32() + off, nxt, half, s);
    std::swap(cur, nxt);
    off += half;
    cur_n = half;
  }
  if (inverse) {
    batch_inverse(line.u32(), line.u32(), off, s);
    batch_inverse(circle.u32(), circle.u32(), n / 2, s);
  }
  VX_HIP(hipStreamSynchronize(s));
  ref.line = line.u32();
  ref.circle = circle.u32();
  ref.log_n = L;
}

// Fold twiddles (src/fri.rs:112-160): layer 0 = circle fold on the eval
// domain (inverse y), layer k >= 1 = line fold on half_odds(le - k)
// (inverse x); entry i uses the point at bitrev(2i).
ProverTables::ProverTables(uint32_t log_n_, uint32_t blowup_bits, hipStream_t s)
    : log_n(log_n_), le(log_n_ + blowup_bits), n_fri(log_n_ + blowup_bits - 3) {
  trace_itw = std::make_unique<TwiddleSet>(make_coset(HALF_COSET, log_n), true, s);
  eval_tw = std::make_unique<TwiddleSet>(eval_domain(le), false, s);
  bytes += trace_itw->line.bytes + trace_itw->circle.bytes + eval_tw->line.bytes + eval_tw->circle.bytes;
  uint64_t total = 0;
  for (uint32_t k = 0; k < n_fri; k++) {
    fold_tw_off.push_back(total);
    total += 1ull << (le - 1 - k);
  }
  fold_tw = DeviceBuf(4 * total);
  bytes += fold_tw.bytes;
  for (uint32_t k = 0; k < n_fri; k++) {
    coset d = k == 0 ? eval_domain(le) : make_coset(HALF_ODDS, le - k);
    fold_twiddle_sources(d.init, d.step, fold_tw.u32() + fold_tw_off[k], 1ull << d.log_size, d.log_size, k == 0, s);
  }
  batch_inverse(fold_tw.u32(), fold_tw.u32(), total, s);
  VX_HIP(hipStreamSynchronize(s));
}

uint64_t ProverTables::estimate_bytes(uint32_t log_n, uint32_t blowup_bits) {
  const uint64_t n = 1ull << log_n, N = n << blowup_bits;
  // line (n) + circle (n/2) per twiddle set, fold tables sum to ~N
  return 4 * (n + n / 2) + 4 * (N + N / 2) + 4 * N;
}

std::shared_ptr<const ProverTables> ProverTables::get(uint32_t log_n, uint32_t blowup_bits, hipStream_t s) {
  static std::mutex mu;
  static std::map<std::tuple<int, uint32_t, uint32_t>, std::weak_ptr<const ProverTables>> cache;
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(dev, log_n, blowup_bits);
  if (auto sp = cache[key].lock()) return sp;
  auto sp = std::make_shared<const ProverTables>(log_n, blowup_bits, s);
  cache[key] = sp;
  return sp;
}

static void check_air(const AirSpec& air) {
  if (air.kind == AIR_FIBONACCI) return;
  if (air.kind == AIR_BATCH) {
    require(air.n_cols >= 1 && air.n_cols <= 4096 && air.log_sizes.size() == air.n_cols,
            "prover: batch needs 1..4096 columns");
    for (uint32_t c = 0; c < air.n_cols; c++)
      require(air.log_sizes[c] >= 14 && air.log_sizes[c] <= 24 && (c == 0 || air.log_sizes[c] >= air.log_sizes[c - 1]),
              "prover: batch log sizes must be 14..24, sorted ascending");
    return;
  }
  require(air.kind == AIR_D2, "prover: unknown AIR");
  require(air.n_cols >= 1 && air.n_cols <= D2_MAX_COLS && air.n_cons >= 1 && air.n_cons <= D2_MAX_CONS,
          "prover: multi-column AIR needs 1..256 columns and 1..1024 constraints");
}

uint64_t StarkProver::estimate_bytes(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits,
                                     const AirSpec& air) {
  const uint32_t le = log_n + blowup_bits, n_fri = le - 3;
  const bool d2 = air.kind == AIR_D2, batch = air.kind == AIR_BATCH;
  const uint64_t W = d2 ? air.n_cols : 1, G = d2 ? d2_groups(air.n_cols) : 1;
  uint64_t b = 0;
  if (batch) {
    for (uint32_t L : air.log_sizes) b += (8ull << L) + (8ull << (L + blowup_bits));  // in, coef, eval, bit-reversed
    b += (16ull << le) + 16ull * air.n_cols + 32ull * air.n_cols;  // quotient, powers, lifted table
  } else {
    b += (4ull << log_n) * W;            // trace
    b += (4ull << le) * W;               // eval
  }
  if (d2) b += (16ull << le) + 32ull * air.n_cons + 16ull * air.n_cons;  // quotient, table, powers
  for (uint32_t k = 0; k < n_fri; k++) b += 16ull << (le - 1 - k);   // FRI layers
  for (uint32_t t = 0; t < G + 1 + n_fri; t++) {                     // trees, levels >= 4
    uint32_t L = t <= G ? le : le - 1 - (t - G - 1);
    uint32_t lo = L < 4 ? L : 4;
    for (uint32_t l = lo; l <= L; l++) b += 32ull << (L - l);
  }
  b += ProofLayout(log_n, blowup_bits, nq, pow_bits, air).total_words * 4;  // device proof
  b += 64 + 16ull * (n_fri + 2) + 4ull * (nq + 3) + 256 + sizeof(DecomDesc) * (G + 1 + n_fri);
  return b;
}

StarkProver::StarkProver(uint32_t log_n, uint32_t blowup_bits, uint32_t n_queries, uint32_t pow_bits,
                         hipStream_t stream, const AirSpec& air)
    : air_((check_air(air), air)),
      layout_((require(log_n >= 4 && blowup_bits >= 1 && log_n + blowup_bits <= 31, "prover: need log_n >= 4, "
                                                                                       "blowup >= 1, log_n + blowup <= 31"),
               log_n),
              blowup_bits, n_queries, pow_bits, air),
      log_n_(log_n),
      blowup_(blowup_bits),
      nq_(n_queries),
      pow_bits_(pow_bits),
      le_(log_n + blowup_bits),
      n_fri_(log_n + blowup_bits - 3),
      n_groups_(layout_.n_groups),
      stream_(stream) {
  require(n_queries >= 1 && n_queries <= 65536, "prover: n_queries must be 1..65536");
  require(pow_bits <= 32, "prover: pow_bits must be <= 32");
  if (!stream_) {
    VX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  const bool d2 = air_.kind == AIR_D2, batch = air_.kind == AIR_BATCH;
  const uint64_t W = d2 ? air_.n_cols : 1;
  const uint64_t n = 1ull << log_n_, N = 1ull << le_;
  if (batch) {
    require(log_n_ == air_.log_sizes.back(), "prover: batch log_n must be the largest column's");
    uint64_t win = 0, wev = 0;
    for (uint32_t L : air_.log_sizes) {
      win += 1ull << L;
      wev += 1ull << (L + blowup_bits);
    }
    n_in_ = win;
    trace_ = DeviceBuf(4 * win);
    coef_ = DeviceBuf(4 * win);
    eval_ = DeviceBuf(4 * wev);
    evr_ = DeviceBuf(4 * wev);
    quot_ = DeviceBuf(16 * N);
    beta_ = DeviceBuf(16ull * air_.n_cols);
    device_bytes_ += trace_.bytes + coef_.bytes + eval_.bytes + evr_.bytes + quot_.bytes + beta_.bytes;
  } else {
    trace_ = DeviceBuf(4 * n * W);
    eval_ = DeviceBuf(4 * N * W);
    device_bytes_ += trace_.bytes + eval_.bytes;
  }
  if (d2) {
    quot_ = DeviceBuf(16 * N);
    std::vector<uint32_t> cd(2 * air_.n_cons), tab(8 * air_.n_cons, 0);
    d2_coeffs(air_.air_seed, air_.n_cons, cd.data());
    for (uint32_t k = 0; k < air_.n_cons; k++) {
      tab[8 * k + 0] = cd[2 * k];
      tab[8 * k + 1] = cd[2 * k + 1];
      tab[8 * k + 2] = d2_col_a(k, air_.n_cols);
      tab[8 * k + 3] = d2_col_b(k, air_.n_cols);
      tab[8 * k + 4] = d2_col_e(k, air_.n_cols);
    }
    cons_ = DeviceBuf(4 * tab.size());
    VX_HIP(hipMemcpy(cons_.p, tab.data(), 4 * tab.size(), hipMemcpyHostToDevice));
    beta_ = DeviceBuf(32ull * air_.n_cons);
    device_bytes_ += quot_.bytes + cons_.bytes + beta_.bytes;
  }
  for (uint32_t k = 0; k < n_fri_; k++) {
    uint64_t words = 1ull << (le_ - 1 - k);
    layers_.emplace_back(4 * 4 * words);
    layer_col_words_.push_back(words);
    device_bytes_ += layers_.back().bytes;
  }
  for (uint32_t t = 0; t < n_groups_ + 1 + n_fri_; t++) {
    uint32_t L = t <= n_groups_ ? le_ : le_ - 1 - (t - n_groups_ - 1);
    uint32_t lo = L < 4 ? L : 4;
    std::vector<uint64_t> off(MAX_LEVELS, 0);
    uint64_t w = 0;
    for (uint32_t l = lo; l <= L; l++) {
      off[l] = w;
      w += 8ull << (L - l);
    }
    trees_.emplace_back(4 * w);
    tree_log_.push_back(L);
    tree_lo_.push_back(lo);
    tree_off_.push_back(off);
    device_bytes_ += trees_.back().bytes;
  }
  // channel (64 B) | alphas | queries | prefix (64 B) | nonce (64 B) | inputs (64 B)
  const size_t alpha_words = 4ull * (n_fri_ + 2);
  const size_t query_words = (nq_ + 3) / 4 * 4;  // keeps the prefix/nonce slots 16-byte aligned
  const size_t small_bytes = 64 + 4 * alpha_words + 4 * query_words + 64 + 64 + 64;
  small_ = DeviceBuf(small_bytes);
  auto* base = static_cast<uint8_t*>(small_.p);
  d_channel_ = reinterpret_cast<Channel*>(base);
  d_alphas_ = reinterpret_cast<uint32_t*>(base + 64);
  d_queries_ = d_alphas_ + alpha_words;
  d_prefix_ = reinterpret_cast<uint32_t*>(base + 64 + 4 * alpha_words + 4 * query_words);
  d_nonce_ = reinterpret_cast<uint64_t*>(reinterpret_cast<uint8_t*>(d_prefix_) + 64);
  d_inputs_ = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(d_prefix_) + 128);
  proof_ = DeviceBuf(layout_.total_words * 4);
  for (int k = 0; k < 2; k++) {
    VX_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_proof_[k]), layout_.total_words * 4, hipHostMallocDefault));
    VX_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&host_proof_dev_[k]), host_proof_[k], 0));
  }
  device_bytes_ += small_.bytes + proof_.bytes;
  for (auto& e : ev_) VX_HIP(hipEventCreate(&e));
  tab_ = ProverTables::get(log_n_, blowup_, stream_);
  if (batch) build_batch();
  build_descs();
}

// Work lists, lifted column table, column offsets and the constant header
// words of the batched-columns proof.
void StarkProver::build_batch() {
  const uint32_t C = air_.n_cols;
  std::vector<LdeBatchCol> cols(C);
  std::vector<uint64_t> off(C + 1, 0);
  std::vector<const uint32_t*> src(C);
  std::vector<uint32_t*> dst(C);
  std::vector<uint32_t> elog(C);
  std::vector<LiftCol> lift(C);
  uint64_t oin = 0, oev = 0;
  for (uint32_t c = 0; c < C; c++) {
    const uint32_t L = air_.log_sizes[c], E = L + blowup_;
    if (c == 0 || L != air_.log_sizes[c - 1]) {
      col_tables_.push_back(L == log_n_ ? tab_ : ProverTables::get(L, blowup_, stream_));
      if (L != log_n_) device_bytes_ += col_tables_.back()->bytes;
    }
    const ProverTables& t = *col_tables_.back();
    cols[c] = {L, trace_.u32() + oin, coef_.u32() + oin, eval_.u32() + oev, t.trace_itw->ref, t.eval_tw->ref};
    off[c] = oin;
    src[c] = eval_.u32() + oev;
    dst[c] = evr_.u32() + oev;
    elog[c] = E;
    lift[c] = {evr_.u32() + oev, E, 0};
    oin += 1ull << L;
    oev += 1ull << E;
  }
  off[C] = oin;
  std::vector<uint8_t> jobs;
  lde_batch_build(cols.data(), C, blowup_, jobs, lde_st_);
  lde_jobs_ = DeviceBuf(jobs.size());
  VX_HIP(hipMemcpy(lde_jobs_.p, jobs.data(), jobs.size(), hipMemcpyHostToDevice));
  bitrev_batch_build(src.data(), dst.data(), elog.data(), C, jobs, br_count_, br_tiles_);
  br_jobs_ = DeviceBuf(jobs.size());
  VX_HIP(hipMemcpy(br_jobs_.p, jobs.data(), jobs.size(), hipMemcpyHostToDevice));
  lift_ = DeviceBuf(sizeof(LiftCol) * C);
  VX_HIP(hipMemcpy(lift_.p, lift.data(), sizeof(LiftCol) * C, hipMemcpyHostToDevice));
  uint32_t sd[8];
  b2_hash_long(air_.log_sizes.data(), 4 * C, sd);
  size_digest_ = DeviceBuf(32);
  VX_HIP(hipMemcpy(size_digest_.p, sd, 32, hipMemcpyHostToDevice));
  fill_off_ = DeviceBuf(8ull * (C + 1));
  VX_HIP(hipMemcpy(fill_off_.p, off.data(), 8ull * (C + 1), hipMemcpyHostToDevice));
  device_bytes_ += lde_jobs_.bytes + br_jobs_.bytes + lift_.bytes + fill_off_.bytes;
  // header words that never change: magic, version, C, blowup, pow, nonce
  // (zero unless the query step writes it), C log sizes, layer counts
  std::vector<uint32_t> hdr(layout_.header_words, 0);
  hdr[0] = KRAK_MAGIC;
  hdr[1] = 4;
  hdr[2] = C;
  hdr[3] = blowup_;
  hdr[4] = pow_bits_;
  for (uint32_t c = 0; c < C; c++) hdr[9 + c] = air_.log_sizes[c];
  VX_HIP(hipMemcpy(proof_.p, hdr.data(), 4 * hdr.size(), hipMemcpyHostToDevice));
  VX_HIP(hipMemcpy(proof_.u32() + layout_.w_nfri, &n_fri_, 4, hipMemcpyHostToDevice));
  VX_HIP(hipMemcpy(proof_.u32() + layout_.w_nfdec, &n_fri_, 4, hipMemcpyHostToDevice));
}

StarkProver::~StarkProver() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_) (void)hipGraphExecDestroy(graph_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& h : host_proof_)
    if (h) (void)hipHostFree(h);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

TreeLevels StarkProver::tree_levels(uint32_t t) const {
  TreeLevels T{};
  T.base = trees_[t].u32();
  T.log_n = tree_log_[t];
  T.lo = tree_lo_[t];
  for (int l = 0; l < MAX_LEVELS; l++) T.off[l] = tree_off_[t][l];
  return T;
}

void StarkProver::build_descs() {
  const bool d2 = air_.kind == AIR_D2;
  const uint64_t N = 1ull << le_;
  const uint32_t G = n_groups_;
  std::vector<DecomDesc> d(G + 1 + n_fri_);
  for (uint32_t t = 0; t < G + 1 + n_fri_; t++) {
    DecomDesc& x = d[t];
    std::memset(&x, 0, sizeof x);
    x.log_n = tree_log_[t];
    x.lo = tree_lo_[t];
    x.tree = trees_[t].u32();
    for (int l = 0; l < MAX_LEVELS; l++) x.off[l] = tree_off_[t][l];
    if (t < G && air_.kind == AIR_BATCH) {  // the lifted tree
      x.kind = DECOM_LIFTED;
      x.n_cols = air_.n_cols;
      x.lift = static_cast<const LiftCol*>(lift_.p);
      x.pair = PAIR_SIBLING;
      x.shift = 0;
      x.proof_word = layout_.w_tdec[t];
    } else if (t < G) {  // trace group t
      x.kind = DECOM_COLS;
      x.n_cols = d2 ? d2_group_cols(air_.n_cols, t) : 1;
      for (uint32_t c = 0; c < x.n_cols; c++) x.col[c] = eval_.u32() + (D2_GROUP * t + c) * N;
      x.pair = d2 ? PAIR_NEXT : PAIR_SIBLING;
      x.shift = 0;
      x.proof_word = layout_.w_tdec[t];
    } else if (t == G) {
      if (d2 || air_.kind == AIR_BATCH) {
        x.kind = DECOM_SOA4;
        for (int c = 0; c < 4; c++) x.col[c] = quot_.u32() + c * N;
      } else {
        x.kind = DECOM_QUOTIENT;
        x.col[0] = eval_.u32();
        x.alpha = d_alphas_;
      }
      x.shift = 0;
      x.proof_word = layout_.w_qdec;
    } else {
      uint32_t k = t - G - 1;
      x.kind = DECOM_SOA4;
      for (int c = 0; c < 4; c++) x.col[c] = layers_[k].u32() + c * layer_col_words_[k];
      x.shift = k + 1;
      x.proof_word = layout_.w_fdec[k];
    }
  }
  descs_ = DeviceBuf(sizeof(DecomDesc) * d.size());
  VX_HIP(hipMemcpy(descs_.p, d.data(), sizeof(DecomDesc) * d.size(), hipMemcpyHostToDevice));
  device_bytes_ += descs_.bytes;
}

// The whole proof after the public inputs are set: kernels only (no host
// synchronisation, no memset), so it can be captured into a graph.
void StarkProver::record(bool timed, hipStream_t s) {
  const ProverTables& tb = *tab_;
  const ProofLayout& L = layout_;
  uint32_t* proof = proof_.u32();
  const uint64_t n = 1ull << log_n_, N = 1ull << le_;
  auto mark = [&](int i) {
    if (timed) VX_HIP(hipEventRecord(ev_[i], s));
  };
  auto alpha = [&](uint32_t i) { return d_alphas_ + 4 * i; };
  auto layer_col = [&](uint32_t k, int c) { return layers_[k].u32() + c * layer_col_words_[k]; };

  const bool d2 = air_.kind == AIR_D2, batch = air_.kind == AIR_BATCH;
  const uint32_t G = n_groups_;
  const uint32_t QT = G;  // quotient tree index; FRI layer k is tree G + 1 + k
  mark(0);
  if (batch) {
    // the header is constant (written at construction) apart from the seed
    // (set_inputs) and the nonce (query step)
    d2_fill_columns(d_inputs_, trace_.u32(), reinterpret_cast<const uint64_t*>(fill_off_.p), air_.n_cols, n_in_, s);
    mark(1);
    mark(2);
    lde_batch_launch(static_cast<const uint8_t*>(lde_jobs_.p), lde_st_, s);
    bitrev_batch_launch(static_cast<const uint8_t*>(br_jobs_.p), br_count_, br_tiles_, s);
    mark(3);
    LeafSrc lt{};
    lt.kind = LEAF_LIFTED;
    lt.lift = static_cast<const LiftCol*>(lift_.p);
    lt.lift_n = air_.n_cols;
    lt.lift_log = le_;
    channel_step(d_channel_, true, size_digest_.u32(), nullptr, nullptr, s);  // sizes first
    merkle_commit(lt, tree_levels(0), tree_lo_[0], RootHook{d_channel_, proof + L.w_troot[0], alpha(0), 0u}, s,
                  timed ? ev_[10] : nullptr, timed ? ev_[11] : nullptr);
    mark(4);
    d2_powers(alpha(0), beta_.u32(), air_.n_cols, s);
    LeafSrc lq{};
    lq.kind = LEAF_LIFTED_RLC;
    lq.lift = lt.lift;
    lq.lift_n = air_.n_cols;
    lq.lift_log = le_;
    lq.beta = beta_.u32();
    for (int c = 0; c < 4; c++) lq.out[c] = quot_.u32() + c * N;
    merkle_commit(lq, tree_levels(QT), tree_lo_[QT], RootHook{d_channel_, proof + L.w_qroot, alpha(1), 0}, s);
    mark(5);
  }
  std::vector<uint32_t> hw;
  if (batch) {
  } else if (!d2) {  // words 3, 4 (a, b) are written by set_inputs
    hw = {0, KRAK_MAGIC, 1, pow_bits_ ? 2u : 1u, 2, log_n_};
    if (pow_bits_) hw.insert(hw.end(), {5, blowup_, 6, pow_bits_});
  } else {    // words 9, 10 (seed) by set_inputs, 7, 8 (nonce) by the query step or stay 0
    hw = {0, KRAK_MAGIC, 1, 3, 2, log_n_, 3, air_.n_cols, 4, air_.n_cons, 5, blowup_, 6, pow_bits_,
          11, (uint32_t)air_.air_seed, 12, (uint32_t)(air_.air_seed >> 32), 13, G};
    if (!pow_bits_) hw.insert(hw.end(), {7, 0, 8, 0});
  }
  if (!batch) {
    hw.insert(hw.end(), {(uint32_t)L.w_nfri, n_fri_, (uint32_t)L.w_nfdec, n_fri_});
    write_header(proof, hw.data(), (uint32_t)hw.size() / 2, s);
  }
  if (batch) {
  } else if (!d2) {
    fib_trace_dev(d_inputs_, trace_.u32(), log_n_, s);
    mark(1);
    ntt_inverse(trace_.u32(), tb.trace_itw->ref, s);
    mark(2);
    ntt_forward(eval_.u32(), trace_.u32(), n - 1, tb.eval_tw->ref, log_n_, s);
  } else {
    const uint32_t W = air_.n_cols;
    d2_trace(d_inputs_, trace_.u32(), W, log_n_, n, s);
    mark(1);
    ntt_inverse(trace_.u32(), tb.trace_itw->ref, s, W, n);
    mark(2);
    ntt_forward(eval_.u32(), trace_.u32(), n - 1, tb.eval_tw->ref, log_n_, s, W, N, n);
  }
  if (!batch) mark(3);

  // trace commitment: one tree per group of <= 16 columns; the channel mixes
  // every root and draws the quotient challenge after the last
  for (uint32_t g = 0; g < G && !batch; g++) {
    LeafSrc lt{};
    lt.kind = LEAF_COLS;
    lt.n_cols = d2 ? d2_group_cols(air_.n_cols, g) : 1;
    for (uint32_t c = 0; c < lt.n_cols; c++) lt.col[c] = eval_.u32() + (D2_GROUP * g + c) * N;
    const bool last = g + 1 == G;
    merkle_commit(lt, tree_levels(g), tree_lo_[g],
                  RootHook{d_channel_, proof + L.w_troot[g], last ? alpha(0) : nullptr, g == 0 ? 1u : 0u}, s,
                  timed && g == 0 ? ev_[10] : nullptr, timed && g == 0 ? ev_[11] : nullptr);
  }
  if (!batch) mark(4);

  if (batch) {
  } else if (!d2) {
    LeafSrc lq{};
    lq.kind = LEAF_QUOTIENT;
    lq.col[0] = eval_.u32();
    lq.alpha_q = alpha(0);
    lq.src_mask = N - 1;
    merkle_commit(lq, tree_levels(QT), tree_lo_[QT], RootHook{d_channel_, proof + L.w_qroot, alpha(1), 0}, s);
  } else {
    d2_weights(alpha(0), cons_.u32(), beta_.u32(), air_.n_cons, s);
    LeafSrc lq{};
    lq.kind = LEAF_D2_QUOTIENT;
    lq.col[0] = eval_.u32();
    lq.n_cols = air_.n_cols;
    lq.col_stride = N;
    lq.n_cons = air_.n_cons;
    lq.cons = cons_.u32();
    lq.beta = beta_.u32();
    lq.src_mask = N - 1;
    for (int c = 0; c < 4; c++) lq.out[c] = quot_.u32() + c * N;
    merkle_commit(lq, tree_levels(QT), tree_lo_[QT], RootHook{d_channel_, proof + L.w_qroot, alpha(1), 0}, s);
  }
  if (!batch) mark(5);

  LeafSrc lc{};
  if (!d2 && !batch) {  // circle fold of the quotient, recomputed from the eval column
    lc.kind = LEAF_FOLD_CIRCLE_Q;
    lc.col[0] = eval_.u32();
    lc.alpha_q = alpha(0);
    lc.src_mask = N - 1;
  } else {    // circle fold of the stored quotient: same pair formula as a line fold
    lc.kind = LEAF_FOLD_LINE;
    for (int c = 0; c < 4; c++) lc.col[c] = quot_.u32() + c * N;
  }
  lc.alpha_f = alpha(1);
  lc.tw = tb.fold_tw.u32() + tb.fold_tw_off[0];
  for (int c = 0; c < 4; c++) lc.out[c] = layer_col(0, c);
  merkle_commit(lc, tree_levels(QT + 1), tree_lo_[QT + 1],
                RootHook{d_channel_, proof + L.w_froot0, n_fri_ > 1 ? alpha(2) : nullptr, 0}, s);
  // Line folds: big layers through the subtree kernels, the last layers of
  // <= 2^TAIL_MAX_LOG values in one k_fri_tail launch.
  uint32_t k = 1;
  for (; k < n_fri_ && le_ - 1 - k > TAIL_MAX_LOG; k++) {
    LeafSrc ll{};
    ll.kind = LEAF_FOLD_LINE;
    for (int c = 0; c < 4; c++) {
      ll.col[c] = layer_col(k - 1, c);
      ll.out[c] = layer_col(k, c);
    }
    ll.tw = tb.fold_tw.u32() + tb.fold_tw_off[k];
    ll.alpha_f = alpha(k + 1);
    merkle_commit(ll, tree_levels(QT + 1 + k), tree_lo_[QT + 1 + k],
                  RootHook{d_channel_, proof + L.w_froot0 + 8ull * k, k + 1 < n_fri_ ? alpha(k + 2) : nullptr, 0}, s);
  }
  if (k < n_fri_) {
    FriTailArgs ta{};
    for (int c = 0; c < 4; c++) ta.in[c] = layer_col(k - 1, c);
    ta.ch = d_channel_;
    ta.n_layers = n_fri_ - k;
    for (uint32_t j = k; j < n_fri_; j++) {
      FriTailLayer& T = ta.layer[j - k];
      for (int c = 0; c < 4; c++) T.out[c] = layer_col(j, c);
      T.tw = tb.fold_tw.u32() + tb.fold_tw_off[j];
      T.tree = trees_[QT + 1 + j].u32();
      T.log_n = le_ - 1 - j;
      T.lo = tree_lo_[QT + 1 + j];
      for (uint32_t l = 0; l <= T.log_n; l++) T.off[l] = (uint32_t)tree_off_[QT + 1 + j][l];
      T.root_out = proof + L.w_froot0 + 8ull * j;
      T.alpha = alpha(j + 1);
      T.next_alpha = j + 1 < n_fri_ ? alpha(j + 2) : nullptr;
    }
    fri_tail(ta, s);
  }
  mark(6);

  QueryArgs q{};
  q.ch = d_channel_;
  for (int c = 0; c < 4; c++) q.last[c] = layer_col(n_fri_ - 1, c);
  q.pow_bits = pow_bits_;
  q.nonce = d_nonce_;
  q.nq = nq_;
  q.domain = N;
  q.queries = d_queries_;
  q.proof = proof;
  q.last_word = L.w_last;
  q.query_word = L.w_nq;
  q.nonce_word = pow_bits_ ? L.w_nonce : 0;
  channel_mix_last(q, d_prefix_, s);
  if (pow_bits_) grind_pow(d_prefix_, pow_bits_, d_nonce_, s);
  channel_draw_queries(q, s);
  mark(7);
  decommit(reinterpret_cast<const DecomDesc*>(descs_.p), G + 1 + n_fri_, d_queries_, nq_, proof, s);
  mark(8);
}

// Captured on a private stream, so work that other threads put on the
// prover's stream meanwhile cannot end up in the graph.
void StarkProver::capture() {
  hipStream_t cs = nullptr;
  VX_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    try {
      record(false, cs);
    } catch (...) {
      if (hipStreamEndCapture(cs, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      (void)hipStreamDestroy(cs);
      throw;
    }
    e = hipStreamEndCapture(cs, &g);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  VX_HIP(e);
}

void StarkProver::tune() {
  if (tuned_) return;
  set_inputs(d_inputs_, proof_.u32(), air_.kind == AIR_D2 ? 9 : (air_.kind == AIR_BATCH ? 7 : 3), 1, 1, stream_);
  record(false, stream_);
  VX_HIP(hipStreamSynchronize(stream_));
  tuned_ = true;
}

void StarkProver::enqueue(uint32_t a, uint32_t b, bool timed) {
  const bool seeded = air_.kind != AIR_FIBONACCI;  // (a, b) = the trace seed's words
  require(seeded || (a < P && b < P), "prove: public inputs must be canonical M31");
  timed_ = timed;
  set_inputs(d_inputs_, proof_.u32(), air_.kind == AIR_D2 ? 9 : (air_.kind == AIR_BATCH ? 7 : 3), a, b, stream_);
  if (timed || !use_graph_) {
    record(timed, stream_);
  } else {
    if (!graph_) {
      tune();
      set_inputs(d_inputs_, proof_.u32(), air_.kind == AIR_D2 ? 9 : (air_.kind == AIR_BATCH ? 7 : 3), a, b,
                 stream_);
      capture();
    }
    VX_HIP(hipGraphLaunch(graph_, stream_));
  }
  cur_ ^= 1;
  copy_to_host(host_proof_dev_[cur_], proof_.u32(), layout_.total_words, stream_);
  if (timed) VX_HIP(hipEventRecord(ev_[9], stream_));
}

const uint8_t* StarkProver::wait(size_t* len) {
  VX_HIP(hipStreamSynchronize(stream_));
  if (len) *len = layout_.total_words * 4;
  return host_proof_[cur_];
}

StageTimes StarkProver::last_times() {
  StageTimes t{};
  if (!timed_) return t;
  VX_HIP(hipEventSynchronize(ev_[9]));
  auto el = [&](int a, int b) {
    float ms = 0;
    VX_HIP(hipEventElapsedTime(&ms, ev_[a], ev_[b]));
    return ms;
  };
  t.trace_ms = el(0, 1);
  t.interpolate_ms = el(1, 2);
  t.lde_ms = el(2, 3);
  t.trace_commit_ms = el(3, 4);
  t.quotient_commit_ms = el(4, 5);
  t.fri_ms = el(5, 6);
  t.queries_ms = el(6, 7);
  t.decommit_ms = el(7, 8);
  t.download_ms = el(8, 9);
  t.total_ms = el(0, 9);
  t.trace_subtree_ms = el(10, 11);
  return t;
}

}  // namespace vx
