// stwo wire-format export of KRAK proofs (src/stwo_export.rs) and the stwo
// Merkle-witness helpers. Host code.
//
// A KRAK decommitment carries, per query, the opened values at the query
// and at its partner row, each with its full auth path. stwo instead keeps
// one `hash_witness` per tree: the sibling hashes, level by level from the
// leaves, of positions whose sibling is not itself known
// (src/stwo_export.rs:262-320); an FRI layer adds `fri_witness`, the values at
// siblings of queried positions that are not queried themselves (:174-206).
// The reference exports its Cairo proofs that way (cairo_proof_to_stwo,
// :401-636); the same conventions are applied here to the KRAK proofs of
// this path (Fibonacci v1/v2, the multi-column AIR v3, the batched columns
// v4). KRAK has no out-of-domain sampling, so `sampled_values` is empty per
// tree; the quotient tree is the FRI first layer.
#include "export.h"

#include <cstring>
#include <map>
#include <set>

#include "vx_blake2s.h"
#include "vx_circle.h"
#include "vx_d2.h"
#include "vx_error.h"

namespace vx {

std::vector<std::array<uint32_t, 8>> hash_witness(const std::vector<uint64_t>& positions,
                                                  const std::vector<const uint32_t*>& paths,
                                                  const std::vector<size_t>& path_words, uint32_t height) {
  std::vector<std::array<uint32_t, 8>> out;
  if (positions.empty() || height == 0) return out;
  require(paths.size() == positions.size() && path_words.size() == positions.size(),
          "hash witness: one auth path per position");
  std::map<uint64_t, size_t> cur;  // position at this level -> auth path (first wins)
  for (size_t i = 0; i < positions.size(); i++) cur.emplace(positions[i], i);
  for (uint32_t level = 0; level < height; level++) {
    for (auto it = cur.begin(); it != cur.end();) {
      auto nx = std::next(it);
      if (nx != cur.end() && nx->first == (it->first ^ 1)) {
        it = std::next(nx);  // both children known
        continue;
      }
      const size_t q = it->second;
      if (8ull * level + 8 <= path_words[q]) {  // a shorter path contributes nothing (as the reference)
        std::array<uint32_t, 8> h;
        std::memcpy(h.data(), paths[q] + 8ull * level, 32);
        out.push_back(h);
      }
      ++it;
    }
    std::map<uint64_t, size_t> up;
    for (auto& kv : cur) up.emplace(kv.first >> 1, kv.second);
    cur.swap(up);
  }
  return out;
}

std::vector<std::array<uint32_t, 4>> fri_witness(const std::vector<uint64_t>& positions, const uint32_t* sib_values) {
  std::map<uint64_t, const uint32_t*> sib;  // first occurrence wins
  for (size_t i = 0; i < positions.size(); i++) sib.emplace(positions[i], sib_values + 4 * i);
  std::vector<std::array<uint32_t, 4>> out;
  for (auto& kv : sib)
    if (!sib.count(kv.first ^ 1)) {
      std::array<uint32_t, 4> v;
      std::memcpy(v.data(), kv.second, 16);
      out.push_back(v);
    }
  return out;
}

std::string reconstruct_root(uint32_t log_height, const std::vector<uint64_t>& positions, const uint32_t* col_values,
                             uint32_t n_cols, const uint32_t* witness, size_t n_witness, uint32_t root[8]) {
  if (positions.empty()) return "no query positions";
  const size_t nq = positions.size();
  std::map<uint64_t, std::array<uint32_t, 8>> layer;
  std::vector<uint32_t> row(n_cols);
  for (size_t q = 0; q < nq; q++) {
    if (layer.count(positions[q])) continue;  // first occurrence wins
    for (uint32_t c = 0; c < n_cols; c++) row[c] = col_values[c * nq + q];
    std::array<uint32_t, 8> h;
    b2_hash_long(row.data(), 4 * n_cols, h.data());
    layer.emplace(positions[q], h);
  }
  size_t wi = 0;
  for (uint32_t level = 0; level < log_height; level++) {
    std::map<uint64_t, std::array<uint32_t, 8>> up;
    for (auto it = layer.begin(); it != layer.end();) {
      const uint64_t pos = it->first;
      const uint32_t *l, *r;
      auto nx = std::next(it);
      if (nx != layer.end() && nx->first == (pos ^ 1)) {
        l = it->second.data();
        r = nx->second.data();
        it = std::next(nx);
      } else {
        if (wi >= n_witness) return "hash_witness too short";
        const uint32_t* w = witness + 8 * wi++;
        l = (pos & 1) ? w : it->second.data();
        r = (pos & 1) ? it->second.data() : w;
        ++it;
      }
      std::array<uint32_t, 8> parent;
      b2_hash_node(l, r, parent.data());
      up.emplace(pos >> 1, parent);
    }
    layer.swap(up);
  }
  if (wi != n_witness) return "hash_witness too long (not fully consumed)";
  if (layer.size() != 1) return "expected exactly one root node";
  std::memcpy(root, layer.begin()->second.data(), 32);
  return "";
}

std::vector<std::array<uint32_t, 4>> last_layer_coeffs(const uint32_t* evals, uint32_t n) {
  require(n >= 2 && n <= 4096 && (n & (n - 1)) == 0, "last layer: size must be a power of two in [2, 4096]");
  uint32_t log_n = 0;
  while ((1u << log_n) < n) log_n++;
  std::vector<qm31> v(n);
  for (uint32_t i = 0; i < n; i++)  // bit-reversed evaluations -> natural order
    std::memcpy(v[bitrev64(i, log_n)].v, evals + 4 * i, 16);
  // inverse line FFT over half_odds(k), half_odds(k - 1), ...: the second
  // half of a chunk holds the evaluations at the negated x of the first
  for (uint32_t chunk = n, dom = log_n; chunk > 1; chunk >>= 1, dom--) {
    const coset c = make_coset(HALF_ODDS, dom);
    const uint32_t half = chunk / 2;
    std::vector<uint32_t> xinv(half);
    for (uint32_t i = 0; i < half; i++) xinv[i] = m31_inv(coset_at(c, i).x);
    for (uint32_t base = 0; base < n; base += chunk)
      for (uint32_t i = 0; i < half; i++) {
        const qm31 a = v[base + i], b = v[base + half + i];
        v[base + i] = qm_add(a, b);
        v[base + half + i] = qm_scale(qm_sub(a, b), xinv[i]);
      }
  }
  const uint32_t ninv = m31_inv(n);
  std::vector<std::array<uint32_t, 4>> out(n);
  // the butterflies leave the coefficients bit-reversed, which is the order
  // stwo's LinePoly stores them in
  for (uint32_t i = 0; i < n; i++) {
    const qm31 s = qm_scale(v[i], ninv);
    std::memcpy(out[i].data(), s.v, 16);
  }
  return out;
}

namespace {

struct Rd {
  const uint8_t* p;
  size_t len, pos = 0;
  uint32_t u32() {
    if (pos + 4 > len) throw Error(ERR_ARG, "stwo export: truncated proof");
    uint32_t v;
    std::memcpy(&v, p + pos, 4);
    pos += 4;
    return v;
  }
};

// One KRAK decommitment: per query, the values and auth path at pos[i] and
// at the partner position spos[i] (pos ^ 1, or the next row for v3 traces).
struct Opened {
  uint32_t width = 0, log_size = 0;
  uint32_t root[8];
  std::string name;
  std::vector<uint64_t> pos, spos;
  std::vector<uint32_t> vals, sibs;
  std::vector<std::vector<uint32_t>> paths, spaths;
};

void read_opened(Rd& r, Opened& o, uint32_t width) {
  o.width = width;
  const uint32_t n = r.u32();
  if (n > (1u << 20) || (uint64_t)n * width * 8 > r.len) throw Error(ERR_ARG, "stwo export: bad decommitment");
  o.vals.resize((size_t)n * width);
  o.sibs.resize((size_t)n * width);
  for (auto& v : o.vals) v = r.u32();
  for (auto& v : o.sibs) v = r.u32();
  for (auto* ps : {&o.paths, &o.spaths}) {
    ps->resize(n);
    for (auto& p : *ps) {
      const uint32_t dep = r.u32();
      if (dep > 64) throw Error(ERR_ARG, "stwo export: bad auth path depth");
      p.resize(8ull * dep);
      for (auto& w : p) w = r.u32();
    }
  }
}

struct Parsed {
  uint32_t blowup = 0, pow_bits = 0, le = 0;
  uint64_t nonce = 0;
  std::vector<Opened> trace;
  Opened quot;
  std::vector<Opened> fri;
  std::vector<uint32_t> last;
  std::vector<uint64_t> queries;
};

// KRAK layouts: src/bin/stark_cli.rs:1051-1230 for v1/v2; v3 and v4 are
// this library's multi-column and batched-columns proofs (read by vx_verify).
Parsed parse(const uint8_t* proof, size_t len, uint32_t blowup_bits) {
  Rd r{proof, len};
  Parsed k;
  if (r.u32() != 0x4B52414Bu) throw Error(ERR_ARG, "stwo export: not a KRAK proof");
  const uint32_t version = r.u32();
  uint32_t log_n = 0;
  std::vector<uint32_t> widths;
  std::vector<std::string> names;
  bool next_row = false;
  auto nonce = [&] {
    uint64_t v = r.u32();
    return v | (uint64_t)r.u32() << 32;
  };
  if (version == 1 || version == 2) {
    log_n = r.u32();
    (void)r.u32();  // public inputs a, b
    (void)r.u32();
    k.blowup = blowup_bits;
    if (version == 2) {
      k.blowup = r.u32();
      k.pow_bits = r.u32();
      k.nonce = nonce();
    }
    widths = {1};
    names = {"trace"};
  } else if (version == 3) {
    log_n = r.u32();
    const uint32_t W = r.u32();
    (void)r.u32();  // constraint count
    k.blowup = r.u32();
    k.pow_bits = r.u32();
    k.nonce = nonce();
    for (int i = 0; i < 4; i++) (void)r.u32();  // trace seed, air seed
    const uint32_t G = r.u32();
    if (W < 1 || W > D2_MAX_COLS || G != d2_groups(W)) throw Error(ERR_ARG, "stwo export: bad column groups");
    for (uint32_t g = 0; g < G; g++) {
      widths.push_back(d2_group_cols(W, g));
      names.push_back("trace_group_" + std::to_string(g));
    }
    next_row = true;
  } else if (version == 4) {
    const uint32_t C = r.u32();
    k.blowup = r.u32();
    k.pow_bits = r.u32();
    k.nonce = nonce();
    (void)r.u32();  // trace seed
    (void)r.u32();
    if (C < 1 || C > 4096) throw Error(ERR_ARG, "stwo export: bad column count");
    for (uint32_t c = 0; c < C; c++) log_n = r.u32();  // ascending; the tree is as tall as the last
    widths = {C};
    names = {"lifted_trace"};
  } else {
    throw Error(ERR_ARG, "stwo export: unsupported KRAK version " + std::to_string(version));
  }
  if (log_n < 1 || k.blowup < 1 || log_n + k.blowup > 31) throw Error(ERR_ARG, "stwo export: bad log sizes");
  k.le = log_n + k.blowup;
  k.trace.resize(widths.size());
  for (auto& t : k.trace)
    for (auto& w : t.root) w = r.u32();
  for (auto& w : k.quot.root) w = r.u32();
  const uint32_t n_fri = r.u32();
  if (n_fri >= k.le) throw Error(ERR_ARG, "stwo export: bad FRI layer count");
  k.fri.resize(n_fri);
  for (auto& f : k.fri)
    for (auto& w : f.root) w = r.u32();
  const uint32_t n_last = r.u32();
  if (n_last < 2 || n_last > 4096 || (n_last & (n_last - 1))) throw Error(ERR_ARG, "stwo export: bad last layer");
  k.last.resize(4ull * n_last);
  for (auto& w : k.last) w = r.u32();
  const uint32_t nq = r.u32();
  if (nq < 1 || nq > (1u << 20)) throw Error(ERR_ARG, "stwo export: bad query count");
  k.queries.resize(nq);
  for (auto& q : k.queries) {
    q = r.u32();
    if (q >> k.le) throw Error(ERR_ARG, "stwo export: query outside the evaluation domain");
  }
  const uint64_t mask = (1ull << k.le) - 1;
  auto place = [&](Opened& o, const std::string& name, uint32_t log_size, uint32_t shift, bool nxt) {
    o.name = name;
    o.log_size = log_size;
    for (uint64_t q : k.queries) {
      o.pos.push_back(q >> shift);
      o.spos.push_back(nxt ? ((q + 1) & mask) : ((q >> shift) ^ 1));
    }
    if (o.vals.size() != (size_t)nq * o.width)
      throw Error(ERR_ARG, "stwo export: decommitment size differs from the query count");
    for (size_t i = 0; i < nq; i++)
      if (o.paths[i].size() != 8ull * log_size || o.spaths[i].size() != 8ull * log_size)
        throw Error(ERR_ARG, "stwo export: auth path length differs from the tree height");
  };
  for (size_t g = 0; g < widths.size(); g++) {
    read_opened(r, k.trace[g], widths[g]);
    place(k.trace[g], names[g], k.le, 0, next_row);
  }
  read_opened(r, k.quot, 4);
  place(k.quot, "quotient", k.le, 0, false);
  if (r.u32() != n_fri) throw Error(ERR_ARG, "stwo export: FRI decommitment count differs from the layer count");
  for (uint32_t j = 0; j < n_fri; j++) {
    read_opened(r, k.fri[j], 4);
    place(k.fri[j], "fri_layer_" + std::to_string(j), k.le - 1 - j, j + 1, false);
  }
  if (r.pos != r.len) throw Error(ERR_ARG, "stwo export: trailing bytes");
  return k;
}

// A set of opened leaves of one tree: sorted unique positions, their values
// (column-major) and the hash witness that proves them.
struct Opening {
  std::vector<uint64_t> positions;
  std::vector<std::vector<uint32_t>> cols;
  std::vector<std::array<uint32_t, 8>> witness;
};

Opening opening(const Opened& o, bool with_partners) {
  using Leaf = std::pair<const uint32_t*, const std::vector<uint32_t>*>;
  std::map<uint64_t, Leaf> leaves;  // first wins
  for (size_t i = 0; i < o.pos.size(); i++) leaves.emplace(o.pos[i], Leaf(o.vals.data() + i * o.width, &o.paths[i]));
  if (with_partners)
    for (size_t i = 0; i < o.pos.size(); i++)
      leaves.emplace(o.spos[i], Leaf(o.sibs.data() + i * o.width, &o.spaths[i]));
  Opening op;
  std::vector<const uint32_t*> paths;
  std::vector<size_t> words;
  op.cols.assign(o.width, {});
  for (auto& kv : leaves) {
    op.positions.push_back(kv.first);
    paths.push_back(kv.second.second->data());
    words.push_back(kv.second.second->size());
    for (uint32_t c = 0; c < o.width; c++) op.cols[c].push_back(kv.second.first[c]);
  }
  op.witness = hash_witness(op.positions, paths, words, o.log_size);
  return op;
}

struct Json {
  std::string s;
  void raw(const char* t) { s += t; }
  void num(uint64_t v) { s += std::to_string(v); }
  void hash(const uint32_t* w) {  // TwoHash: 32 bytes, words little-endian
    s += '[';
    for (int i = 0; i < 32; i++) {
      if (i) s += ',';
      num((w[i / 4] >> (8 * (i % 4))) & 0xFFu);
    }
    s += ']';
  }
  void qm(const uint32_t* v) {
    s += '[';
    for (int i = 0; i < 4; i++) {
      if (i) s += ',';
      num(v[i]);
    }
    s += ']';
  }
  template <class T, class F>
  void list(const T& xs, F f) {
    s += '[';
    bool first = true;
    for (const auto& x : xs) {
      if (!first) s += ',';
      first = false;
      f(x);
    }
    s += ']';
  }
  void hashes(const std::vector<std::array<uint32_t, 8>>& hs) {
    list(hs, [&](const std::array<uint32_t, 8>& h) { hash(h.data()); });
  }
  void opening(const Opening& op, uint32_t log_size) {
    raw("{\"log_size\":");
    num(log_size);
    raw(",\"positions\":");
    list(op.positions, [&](uint64_t p) { num(p); });
    raw(",\"values\":");
    list(op.cols, [&](const std::vector<uint32_t>& col) { list(col, [&](uint32_t v) { num(v); }); });
    raw("}");
  }
};

std::vector<size_t> iota(size_t n) {
  std::vector<size_t> v(n);
  for (size_t i = 0; i < n; i++) v[i] = i;
  return v;
}

}  // namespace

std::string export_stwo_json(const uint8_t* proof, size_t len, uint32_t blowup_bits) {
  require(proof != nullptr, "stwo export: null proof");
  const Parsed k = parse(proof, len, blowup_bits);
  std::vector<const Opened*> trees;
  for (auto& t : k.trace) trees.push_back(&t);
  trees.push_back(&k.quot);
  // decommitments: trace trees at the queries and their partner rows
  // (build_trace_witness, src/stwo_export.rs:645-670), the quotient tree at
  // the queries (:563-567); FRI layers at whole fold pairs (:208-259)
  std::vector<Opening> tree_open;
  for (auto* t : trees) tree_open.push_back(opening(*t, t != &k.quot));
  const Opening first = opening(k.quot, true);
  std::vector<Opening> inner;
  for (auto& f : k.fri) inner.push_back(opening(f, true));
  const uint32_t n_last = (uint32_t)(k.last.size() / 4);
  uint32_t log_last = 0;
  while ((1u << log_last) < n_last) log_last++;
  const size_t nq = k.queries.size();

  Json j;
  j.raw("{\"config\":{\"log_blowup_factor\":");
  j.num(k.blowup);
  j.raw(",\"n_queries\":");
  j.num(nq);
  j.raw(",\"pow_bits\":");
  j.num(k.pow_bits);
  j.raw(",\"log_last_layer_degree_bound\":");
  j.num(log_last);
  j.raw("},\"tree_meta\":");
  j.list(trees, [&](const Opened* t) {
    j.raw("{\"name\":\"");
    j.raw(t->name.c_str());
    j.raw("\",\"n_cols\":");
    j.num(t->width);
    j.raw(",\"log_eval_size\":");
    j.num(t->log_size);
    j.raw("}");
  });
  j.raw(",\"commitments\":");
  j.list(trees, [&](const Opened* t) { j.hash(t->root); });
  j.raw(",\"sampled_values\":");
  j.list(trees, [&](const Opened*) { j.raw("[]"); });
  j.raw(",\"decommitments\":");
  j.list(tree_open, [&](const Opening& op) {
    j.raw("{\"hash_witness\":");
    j.hashes(op.witness);
    j.raw("}");
  });
  j.raw(",\"queried_values\":");  // [tree][col][query], src/stwo_export.rs:498-516
  j.list(trees, [&](const Opened* t) {
    j.list(iota(t->width), [&](size_t c) {
      j.list(iota(nq), [&](size_t q) { j.num(t->vals[q * t->width + c]); });
    });
  });
  j.raw(",\"query_indices\":");
  j.list(k.queries, [&](uint64_t q) { j.num(q); });
  j.raw(",\"proof_of_work\":");
  j.num(k.nonce);
  auto layer = [&](const Opened& o, const Opening& op) {
    j.raw("{\"fri_witness\":");
    j.list(fri_witness(o.pos, o.sibs.data()), [&](const std::array<uint32_t, 4>& v) { j.qm(v.data()); });
    j.raw(",\"decommitment\":{\"hash_witness\":");
    j.hashes(op.witness);
    j.raw("},\"commitment\":");
    j.hash(o.root);
    j.raw("}");
  };
  j.raw(",\"fri_proof\":{\"first_layer\":");
  layer(k.quot, first);
  j.raw(",\"inner_layers\":");
  j.list(iota(k.fri.size()), [&](size_t f) { layer(k.fri[f], inner[f]); });
  j.raw(",\"last_layer_poly\":");
  j.list(last_layer_coeffs(k.last.data(), n_last), [&](const std::array<uint32_t, 4>& v) { j.qm(v.data()); });
  // extension (not part of stwo's proof): the leaves each hash witness
  // proves, so that every root can be rebuilt from this document alone
  j.raw("},\"krak_openings\":{\"trees\":");
  j.list(iota(trees.size()), [&](size_t t) { j.opening(tree_open[t], trees[t]->log_size); });
  j.raw(",\"first_layer\":");
  j.opening(first, k.quot.log_size);
  j.raw(",\"inner_layers\":");
  j.list(iota(k.fri.size()), [&](size_t f) { j.opening(inner[f], k.fri[f].log_size); });
  j.raw(",\"last_layer_evaluations\":");
  j.list(iota(n_last), [&](size_t i) { j.qm(k.last.data() + 4 * i); });
  j.raw("}}");
  return j.s;
}

}  // namespace vx
