// Autotuner of the Merkle subtree schedule (tune.cpp); the lookup used by
// merkle_commit is tuned_shape in vx_internal.h.
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

}  // namespace vx
