// Fibonacci STARK verifier over KRAK proof bytes (src/verifier.rs:39-310,
// deserialisation src/bin/stark_cli.rs:1189-1230). Host code: it replays
// the transcript, re-derives the queries, checks every Merkle path and every
// FRI fold equation. KRAK v2 (PoW) additionally checks the nonce before the
// query draws.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "prover.h"

namespace vx {

namespace {

struct Reader {
  const uint8_t* p;
  size_t len, pos = 0;
  bool bad = false;
  uint32_t u32() {
    if (pos + 4 > len) {
      bad = true;
      return 0;
    }
    uint32_t v;
    std::memcpy(&v, p + pos, 4);
    pos += 4;
    return v;
  }
};

struct Decom {
  uint32_t width = 1;
  std::vector<uint32_t> vals, sibs;  // n * width
  std::vector<std::vector<uint32_t>> paths, spaths;
  size_t n() const { return vals.size() / width; }
};

bool read_decom(Reader& r, Decom& d, uint32_t width) {
  d.width = width;
  uint32_t n = r.u32();
  if (r.bad || n > (1u << 20) || (size_t)n * width * 8 > r.len) return false;
  d.vals.resize((size_t)n * width);
  d.sibs.resize((size_t)n * width);
  for (auto& v : d.vals) v = r.u32();
  for (auto& v : d.sibs) v = r.u32();
  for (auto* ps : {&d.paths, &d.spaths}) {
    ps->resize(n);
    for (uint32_t i = 0; i < n; i++) {
      uint32_t dep = r.u32();
      if (r.bad || dep > 64) return false;
      (*ps)[i].resize(8ull * dep);
      for (auto& w : (*ps)[i]) w = r.u32();
    }
  }
  return !r.bad;
}

void leaf_hash(const uint32_t* w, uint32_t width, uint32_t out[8]) { b2_hash_words(w, width, out); }

// src/merkle.rs:520-566
bool check_path(const uint32_t root[8], const uint32_t leaf[8], uint64_t idx, const std::vector<uint32_t>& path) {
  uint32_t cur[8];
  std::memcpy(cur, leaf, 32);
  for (size_t k = 0; k < path.size() / 8; k++) {
    const uint32_t* sib = path.data() + 8 * k;
    uint32_t nxt[8];
    if (idx % 2 == 0) b2_hash_node(cur, sib, nxt);
    else b2_hash_node(sib, cur, nxt);
    std::memcpy(cur, nxt, 32);
    idx /= 2;
  }
  return std::memcmp(cur, root, 32) == 0;
}

// verifier.rs:274-310
bool check_decom(const Decom& d, const uint32_t root[8], const std::vector<uint64_t>& idx, uint32_t nq) {
  if (d.n() != nq || d.paths.size() != nq) return false;
  for (uint32_t q = 0; q < nq; q++) {
    uint32_t h[8];
    leaf_hash(d.vals.data() + (size_t)q * d.width, d.width, h);
    if (!check_path(root, h, idx[q], d.paths[q])) return false;
    leaf_hash(d.sibs.data() + (size_t)q * d.width, d.width, h);
    if (!check_path(root, h, idx[q] ^ 1, d.spaths[q])) return false;
  }
  return true;
}

// verifier.rs:30-36
uint32_t fold_twiddle_at(const coset& dom, uint64_t folded, bool circle) {
  cpt p = coset_at(dom, bitrev64(folded << 1, dom.log_size));
  return m31_inv(circle ? p.y : p.x);
}

qm31 q_at(const std::vector<uint32_t>& v, size_t q) { return {{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]}}; }

}  // namespace

bool verify_proof(const uint8_t* bytes, size_t len, uint32_t blowup_bits, uint32_t nq, std::string& err) {
  auto fail = [&](const std::string& m) {
    err = m;
    return false;
  };
  err.clear();
  Reader r{bytes, len};
  if (r.u32() != 0x4B52414Bu) return fail("bad magic");
  const uint32_t version = r.u32();
  if (version != 1 && version != 2) return fail("unsupported proof version " + std::to_string(version));
  const uint32_t log_n = r.u32(), a = r.u32(), b = r.u32();
  uint32_t pow_bits = 0;
  uint64_t nonce = 0;
  if (version == 2) {
    uint32_t bb = r.u32();
    pow_bits = r.u32();
    nonce = r.u32();
    nonce |= (uint64_t)r.u32() << 32;
    if (bb != blowup_bits) return fail("blowup mismatch");
    if (pow_bits > 128) return fail("bad pow bits");
  }
  uint32_t troot[8], qroot[8];
  for (auto& w : troot) w = r.u32();
  for (auto& w : qroot) w = r.u32();
  const uint32_t n_fri = r.u32();
  if (r.bad || n_fri > 64) return fail("truncated proof");
  std::vector<uint32_t> froots(8ull * n_fri);
  for (auto& w : froots) w = r.u32();
  const uint32_t n_last = r.u32();
  if (r.bad || n_last > 4096) return fail("truncated proof");
  std::vector<uint32_t> last(4ull * n_last);
  for (auto& w : last) w = r.u32();
  const uint32_t n_q = r.u32();
  if (r.bad || n_q > (1u << 20)) return fail("truncated proof");
  std::vector<uint64_t> qidx(n_q);
  for (auto& v : qidx) v = r.u32();
  Decom td, qd;
  if (!read_decom(r, td, 1) || !read_decom(r, qd, 4)) return fail("truncated proof");
  const uint32_t n_fd = r.u32();
  if (r.bad || n_fd > 64) return fail("truncated proof");
  std::vector<Decom> fd(n_fd);
  for (auto& d : fd)
    if (!read_decom(r, d, 4)) return fail("truncated proof");

  const uint32_t le = log_n + blowup_bits;
  if (log_n < 1 || le > 31) return fail("bad log size");
  const uint64_t N = 1ull << le;
  // verifier.rs:46-53
  if (a >= P || b >= P) return fail("Public inputs out of M31 range");
  if (n_q != nq) return fail("Expected " + std::to_string(nq) + " queries, got " + std::to_string(n_q));
  // verifier.rs:56-70: transcript replay
  Channel ch;
  ch_init(ch);
  ch_mix_digest(ch, troot);
  (void)ch_draw_felt(ch);
  ch_mix_digest(ch, qroot);
  std::vector<qm31> alphas;
  alphas.push_back(ch_draw_felt(ch));
  for (uint32_t k = 0; k < n_fri; k++) {
    ch_mix_digest(ch, froots.data() + 8 * k);
    alphas.push_back(ch_draw_felt(ch));
  }
  // verifier.rs:73-87
  if (n_fri != le - 3) return fail("Expected " + std::to_string(le - 3) + " FRI layers, got " + std::to_string(n_fri));
  if (n_last != 8) return fail("Expected 8 FRI last layer values, got " + std::to_string(n_last));
  // verifier.rs:90-96
  ch_mix_felts(ch, last.data(), n_last);
  if (version == 2) {
    uint32_t prefix[8];
    ch_pow_prefix(ch, pow_bits, prefix);
    if (pow_trailing_zeros(prefix, nonce) < pow_bits) return fail("PoW nonce invalid");
    ch_mix_u64(ch, nonce);
  }
  for (uint32_t i = 0; i < nq; i++)
    if (ch_draw_number(ch, N) != qidx[i]) return fail("Query indices don't match Fiat-Shamir derivation");
  // verifier.rs:99-104
  bool tz = true, qz = true;
  for (int i = 0; i < 8; i++) {
    tz = tz && troot[i] == 0;
    qz = qz && qroot[i] == 0;
  }
  if (tz) return fail("Trace commitment is zero");
  if (qz) return fail("Quotient commitment is zero");
  // verifier.rs:107-144
  if (!check_decom(td, troot, qidx, nq)) return fail("trace auth path invalid");
  if (!check_decom(qd, qroot, qidx, nq)) return fail("quotient auth path invalid");
  std::vector<uint64_t> fidx(nq);
  for (uint32_t i = 0; i < nq; i++) fidx[i] = qidx[i] / 2;
  for (uint32_t k = 0; k < n_fd && k < n_fri; k++) {
    if (!check_decom(fd[k], froots.data() + 8 * k, fidx, nq))
      return fail("FRI layer " + std::to_string(k) + " auth path invalid");
    for (auto& v : fidx) v /= 2;
  }
  // verifier.rs:153-233: fold equations
  if (n_fd == 0) return fail("missing FRI decommitments");
  for (uint32_t k = 0; k < n_fd; k++)
    if (fd[k].n() != nq) return fail("FRI layer decommitment size mismatch");
  for (uint32_t q = 0; q < nq; q++) {
    uint64_t ci = qidx[q];
    uint32_t cl = le;
    qm31 f0 = q_at(qd.vals, q), f1 = q_at(qd.sibs, q);
    if (ci % 2) std::swap(f0, f1);
    qm31 e = fri_fold_pair(f0, f1, alphas[0], fold_twiddle_at(eval_domain(cl), ci / 2, true));
    if (!qm_eq(e, q_at(fd[0].vals, q))) return fail("Circle fold mismatch at query " + std::to_string(q));
    ci /= 2;
    cl--;
    for (uint32_t k = 0; k + 1 < n_fd; k++) {
      qm31 v = q_at(fd[k].vals, q), s = q_at(fd[k].sibs, q);
      if (ci % 2) std::swap(v, s);
      qm31 ex = fri_fold_pair(v, s, alphas[k + 1], fold_twiddle_at(make_coset(HALF_ODDS, cl), ci / 2, false));
      if (!qm_eq(ex, q_at(fd[k + 1].vals, q)))
        return fail("Line fold mismatch at query " + std::to_string(q) + ", layer " + std::to_string(k));
      ci /= 2;
      cl--;
    }
    if (ci >= n_last) return fail("FRI last layer index out of bounds");
    if (!qm_eq(q_at(fd[n_fd - 1].vals, q), q_at(last, ci)))
      return fail("FRI last layer mismatch at query " + std::to_string(q));
  }
  return true;
}

}  // namespace vx
