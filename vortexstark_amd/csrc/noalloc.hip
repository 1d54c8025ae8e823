// Allocation-free variant of the lifted Merkle root (vx_merkle_root_lifted)
// for callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The existing entry point allocates two scratch buffers per call: the table
// of lifted columns the leaf producer reads, and the stored tree levels from
// the subtree tile up (the tile roots and every level above them, which the
// tree-top kernels reduce to the root). Here both live in one file-scope
// device array of NOALLOC_SCRATCH_BYTES. Each GPU holds its own copy of the
// array (at its own address) from the first call on that device until the
// process ends. Calls on one GPU share it, so calls that overlap on one GPU
// (two streams, or two host threads) race and must be serialised by the
// caller. Nothing in the scratch has to start at zero: the table is written
// by the call and every stored tree node is written before it is read.
#include <map>
#include <mutex>

#include "vx_internal.h"

namespace vx {

constexpr uint64_t NOALLOC_SCRATCH_BYTES = 8ull << 20;
constexpr uint32_t NOALLOC_MAX_COLS = 4096;
constexpr uint32_t LIFT_CHUNK = 64;  // columns per table-store launch (1 KiB of kernel arguments)

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

// The device address of `scratch` on the current device. In host code the
// array's name is only a host-side stand-in, so the address is looked up
// once per device with that device current and kept.
static uint8_t* scratch_ptr() {
  static std::mutex mu;
  static std::map<int, uint8_t*> by_device;
  int dev = 0;
  VX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = by_device.find(dev);
  if (it != by_device.end()) return it->second;
  void* p = nullptr;
  VX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  require(p != nullptr, "noalloc: no device address for the scratch array");
  by_device[dev] = static_cast<uint8_t*>(p);
  return static_cast<uint8_t*>(p);
}

struct LiftChunk {
  LiftCol c[LIFT_CHUNK];
  uint32_t n;
};

// Copies one chunk of the column table from the kernel arguments into the
// scratch (no host staging buffer, no allocation).
__global__ __launch_bounds__(LIFT_CHUNK) void k_store_lift_table(LiftChunk ch, LiftCol* dst) {
  const uint32_t t = threadIdx.x;
  if (t < ch.n) dst[t] = ch.c[t];
}

static uint64_t round256(uint64_t b) { return (b + 255) & ~255ull; }

void merkle_root_lifted_noalloc(const uint32_t* const* h_cols, const uint32_t* h_log_sizes, uint32_t n_cols,
                                uint32_t* d_root, hipStream_t s) {
  require(d_root != nullptr && h_cols != nullptr && h_log_sizes != nullptr && n_cols >= 1,
          "lifted_noalloc: null argument");
  require(n_cols <= NOALLOC_MAX_COLS, "lifted_noalloc: at most 4096 columns");
  const uint32_t lg = h_log_sizes[n_cols - 1];
  require(lg >= 4 && lg <= 31, "lifted_noalloc: the largest column needs 16..2^31 rows");
  for (uint32_t c = 0; c < n_cols; c++) {
    require(h_cols[c] && h_log_sizes[c] >= 1 && h_log_sizes[c] <= lg, "lifted_noalloc: column larger than the tree");
    require(c == 0 || h_log_sizes[c] >= h_log_sizes[c - 1], "lifted_noalloc: columns must be sorted by ascending size");
  }
  LeafSrc L{};
  L.kind = LEAF_LIFTED;
  L.lift_n = n_cols;
  L.lift_log = lg;
  // The built-in shape, never the autotuner: tuning times candidate shapes
  // with events and a stream synchronisation.
  const TreeShape shape = merkle_default_shape(L, lg);
  TreeLevels T{};
  uint64_t words = 0;
  T.log_n = lg;
  T.lo = shape.tile_log;
  for (uint32_t l = T.lo; l <= lg; l++) {
    T.off[l] = words;
    words += 8ull << (lg - l);
  }
  const uint64_t table_bytes = round256(sizeof(LiftCol) * n_cols);
  const uint64_t need = table_bytes + 4 * words;
  if (need > NOALLOC_SCRATCH_BYTES)
    throw Error(ERR_ARG, "lifted_noalloc: the column table and tree levels need " + std::to_string(need) +
                             " bytes of scratch, above the 8 MiB maximum (use vx_merkle_root_lifted)");
  uint8_t* base = scratch_ptr();
  LiftCol* table = reinterpret_cast<LiftCol*>(base);
  for (uint32_t c0 = 0; c0 < n_cols; c0 += LIFT_CHUNK) {
    LiftChunk ch{};
    ch.n = n_cols - c0 < LIFT_CHUNK ? n_cols - c0 : LIFT_CHUNK;
    for (uint32_t k = 0; k < ch.n; k++) ch.c[k] = {h_cols[c0 + k], h_log_sizes[c0 + k], 0};
    hipLaunchKernelGGL(k_store_lift_table, dim3(1), dim3(LIFT_CHUNK), 0, s, ch, table + c0);
    VX_HIP(hipGetLastError());
  }
  L.lift = table;
  T.base = reinterpret_cast<uint32_t*>(base + table_bytes);
  RootHook hook{};
  hook.root_out = d_root;
  merkle_commit_shaped(L, T, T.lo, hook, s, shape);
}

}  // namespace vx
