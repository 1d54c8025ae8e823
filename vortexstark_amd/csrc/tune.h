// Autotuner of the Merkle subtree schedule and of the circle NTT's plan
// (tune.cpp); the lookups are tuned_shape and tuned_ntt_plan in
// vx_internal.h.
#pragma once
#include <stdint.h>

#include <vector>

namespace vx {

// enabled = false: every commitment uses the built-in shape. cache_file
// (may be null): choices are loaded from it now and appended as measured.
void tuner_configure(bool enabled, const char* cache_file);

struct TunedEntry {
  uint32_t device, kind, n_cols, n_cons, lift_n, log_n, store_lo, tile_log, lpt;
};
std::vector<TunedEntry> tuner_entries();
// NTT plans chosen so far: (device, fwd, log_n, eff_bits, n_cols, stwo,
// split, order) per entry.
std::vector<uint32_t> tuner_ntt_entries();

}  // namespace vx
