// stwo wire-format export of KRAK proofs and the stwo Merkle-witness
// helpers (src/stwo_export.rs). Host code, no HIP dependency.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace vx {

// src/stwo_export.rs:262-320 (auth_paths_to_hash_witness): positions[i] is
// a leaf, paths[i] its auth path (8 words per level, leaf level first).
std::vector<std::array<uint32_t, 8>> hash_witness(const std::vector<uint64_t>& positions,
                                                  const std::vector<const uint32_t*>& paths,
                                                  const std::vector<size_t>& path_words, uint32_t height);
// src/stwo_export.rs:174-206 (dedup_fri_witness).
std::vector<std::array<uint32_t, 4>> fri_witness(const std::vector<uint64_t>& positions, const uint32_t* sib_values);
// src/stwo_export.rs:697-780 (reconstruct_root_from_witness); col_values is
// column-major, n_cols x positions.size(). Returns "" or the failure.
std::string reconstruct_root(uint32_t log_height, const std::vector<uint64_t>& positions, const uint32_t* col_values,
                             uint32_t n_cols, const uint32_t* witness, size_t n_witness, uint32_t root[8]);
// src/fri.rs:294-335 then the bit-reversal of src/cairo_air/prover.rs:2896-2900:
// last-layer evaluations (bit-reversed order on the half_odds line domain)
// to the bit-reversed LinePoly coefficients stwo's LinePoly::new takes.
std::vector<std::array<uint32_t, 4>> last_layer_coeffs(const uint32_t* evals, uint32_t n);
// The whole KRAK proof (v1..v4) as stwo-shaped JSON.
std::string export_stwo_json(const uint8_t* proof, size_t len, uint32_t blowup_bits);

}  // namespace vx
