// Batched uneven-column commitment (see batch.h).
#include "batch.h"

#include <cstring>

namespace vx {

BatchCommitter::BatchCommitter(const uint32_t* log_sizes, uint32_t n_cols, uint32_t blowup_bits, bool bit_reverse,
                               hipStream_t stream)
    : blowup_(blowup_bits), bitrev_(bit_reverse), stream_(stream) {
  require(log_sizes && n_cols >= 1 && n_cols <= 4096, "batch: 1..4096 columns");
  require(blowup_bits >= 1 && blowup_bits <= 7, "batch: blowup_bits must be 1..7");
  logs_.assign(log_sizes, log_sizes + n_cols);
  for (uint32_t c = 0; c < n_cols; c++) {
    require(logs_[c] >= 14 && logs_[c] <= 24 && logs_[c] + blowup_bits <= 31, "batch: log sizes must be 14..24");
    require(c == 0 || logs_[c] >= logs_[c - 1], "batch: columns must be sorted by ascending log size");
  }
  if (!stream_) {
    VX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  lift_log_ = logs_.back() + blowup_bits;
  uint64_t win = 0, wev = 0;
  for (uint32_t L : logs_) {
    off_in_.push_back(win);
    off_ev_.push_back(wev);
    win += 1ull << L;
    wev += 1ull << (L + blowup_bits);
  }
  in_ = DeviceBuf(4 * win);
  coef_ = DeviceBuf(4 * win);
  eval_ = DeviceBuf(4 * wev);
  if (bitrev_) evr_ = DeviceBuf(4 * wev);
  device_bytes_ += in_.bytes + coef_.bytes + eval_.bytes + evr_.bytes;

  // twiddle tables per distinct size (shared with provers of that size)
  std::vector<LdeBatchCol> cols(n_cols);
  for (uint32_t c = 0; c < n_cols; c++) {
    if (c == 0 || logs_[c] != logs_[c - 1]) {
      tables_.push_back(ProverTables::get(logs_[c], blowup_bits, stream_));
      device_bytes_ += tables_.back()->bytes;
    }
    const ProverTables& t = *tables_.back();
    cols[c] = {logs_[c], in_.u32() + off_in_[c], coef_.u32() + off_in_[c], eval_.u32() + off_ev_[c],
               t.trace_itw->ref, t.eval_tw->ref};
  }
  std::vector<uint8_t> jobs;
  lde_batch_build(cols.data(), n_cols, blowup_bits, jobs, st_);
  jobs_ = DeviceBuf(jobs.size());
  VX_HIP(hipMemcpy(jobs_.p, jobs.data(), jobs.size(), hipMemcpyHostToDevice));

  std::vector<const uint32_t*> src(n_cols);
  std::vector<uint32_t*> dst(n_cols);
  std::vector<uint32_t> elog(n_cols);
  std::vector<LiftCol> lift(n_cols);
  for (uint32_t c = 0; c < n_cols; c++) {
    src[c] = eval_.u32() + off_ev_[c];
    dst[c] = bitrev_ ? evr_.u32() + off_ev_[c] : nullptr;
    elog[c] = logs_[c] + blowup_bits;
    lift[c] = {column(c, 2), elog[c], 0};
  }
  if (bitrev_) {
    std::vector<uint8_t> bj;
    bitrev_batch_build(src.data(), dst.data(), elog.data(), n_cols, bj, br_count_, br_tiles_);
    br_jobs_ = DeviceBuf(bj.size());
    VX_HIP(hipMemcpy(br_jobs_.p, bj.data(), bj.size(), hipMemcpyHostToDevice));
  }
  lift_ = DeviceBuf(sizeof(LiftCol) * n_cols);
  VX_HIP(hipMemcpy(lift_.p, lift.data(), sizeof(LiftCol) * n_cols, hipMemcpyHostToDevice));

  // lifted tree: levels >= 4 stored (as the prover's trees)
  const uint32_t S = lift_log_, lo = 4;
  uint64_t w = 0;
  tl_.log_n = S;
  tl_.lo = lo;
  for (uint32_t l = lo; l <= S; l++) {
    tl_.off[l] = w;
    w += 8ull << (S - l);
  }
  tree_ = DeviceBuf(4 * w);
  tl_.base = tree_.u32();
  root_ = DeviceBuf(64);
  device_bytes_ += jobs_.bytes + br_jobs_.bytes + lift_.bytes + tree_.bytes + root_.bytes;
  for (auto& e : ev_) VX_HIP(hipEventCreate(&e));
}

BatchCommitter::~BatchCommitter() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

uint32_t* BatchCommitter::column(uint32_t c, int which) const {
  require(c < logs_.size(), "batch: column out of range");
  if (which == 0) return in_.u32() + off_in_[c];
  if (which == 1) return coef_.u32() + off_in_[c];
  require(which == 2, "batch: which must be 0, 1 or 2");
  return (bitrev_ ? evr_.u32() : eval_.u32()) + off_ev_[c];
}

void BatchCommitter::enqueue(bool timed) {
  hipStream_t s = stream_;
  timed_ = timed;
  if (timed) VX_HIP(hipEventRecord(ev_[0], s));
  lde_batch_launch(static_cast<const uint8_t*>(jobs_.p), st_, s);
  if (timed) VX_HIP(hipEventRecord(ev_[1], s));
  if (bitrev_) bitrev_batch_launch(static_cast<const uint8_t*>(br_jobs_.p), br_count_, br_tiles_, s);
  if (timed) VX_HIP(hipEventRecord(ev_[2], s));
  LeafSrc L{};
  L.kind = LEAF_LIFTED;
  L.lift = static_cast<const LiftCol*>(lift_.p);
  L.lift_n = n_cols();
  L.lift_log = lift_log_;
  RootHook hook{};
  hook.root_out = root_.u32();
  merkle_commit(L, tl_, tl_.lo, hook, s, nullptr, nullptr);
  if (timed) VX_HIP(hipEventRecord(ev_[3], s));
}

void BatchCommitter::wait(uint32_t root[8]) {
  VX_HIP(hipStreamSynchronize(stream_));
  if (root) VX_HIP(hipMemcpy(root, root_.p, 32, hipMemcpyDeviceToHost));
}

void BatchCommitter::stage_times(float out[4]) {
  for (int i = 0; i < 4; i++) out[i] = 0;
  if (!timed_) return;
  VX_HIP(hipEventSynchronize(ev_[3]));
  VX_HIP(hipEventElapsedTime(&out[0], ev_[0], ev_[1]));
  VX_HIP(hipEventElapsedTime(&out[1], ev_[1], ev_[2]));
  VX_HIP(hipEventElapsedTime(&out[2], ev_[2], ev_[3]));
  VX_HIP(hipEventElapsedTime(&out[3], ev_[0], ev_[3]));
}

}  // namespace vx
