// Twisted circle NTT for gfx950 (interpolate / evaluate / LDE).
//
// Semantics: cuda/circle_ntt.cu:74-117 (butterflies, layer indexing) and
// :248-312 (layer order, 1/n scale). Bit j of the index is butterflied by
// reference layer j (j = 0 is the circle layer, twiddle C[i >> 1]; j >= 1 is
// line layer j - 1, twiddle T_{j-1}[i >> (j+1)]). Evaluate runs bits
// high -> low, interpolate low -> high.
//
// MI355X design: instead of one HBM round trip per layer, a pass transforms
// k consecutive bits [j0, j0 + k) of a 2^k x 2^cbits tile held in LDS
// (64 KiB per workgroup, 16384 words). Rows are 2^j0 apart in HBM; the
// 2^cbits consecutive columns of a row give 64-byte (or longer) contiguous
// runs. The lowest pass is contiguous (j0 = 0, 14 bits), so a 2^24 transform
// takes two passes. The LDE's first log(N/n) evaluate layers only copy the
// zero-padded coefficients, so the first pass reads coefficient i & (n-1)
// directly and those layers are never executed.
#include "vx_internal.h"

namespace vx {

static constexpr uint32_t NTT_THREADS = 512;
static constexpr uint32_t NTT_TILE_LOG = 14;   // words per tile (64 KiB of LDS)
static constexpr uint32_t NTT_STRIDED_MAX = 10; // bits per strided pass (>= 16 columns)

// Kernel variants (template MODE bits):
//   NTT_STWO  bit 0 takes stwo's circle twiddles, derived from line layer 0:
//             butterfly t uses +-L0[(t >> 1) ^ 1], negated when t & 3 is 1 or 2
//             (that is (y, -y, -x, x)[t & 3] of the pair (L0[2p], L0[2p + 1]),
//             p = t >> 2, cuda/circle_ntt_stwo.cu:214-234);
//   NTT_COLS  column blockIdx.y of the batch is cols[blockIdx.y] (in place).
enum : int { NTT_STWO = 1, NTT_COLS = 2 };

struct PassArgs {
  const uint32_t* tw[16];  // twiddle table of bit j0 + b, for local bit b
  uint32_t j0, k, cbits;
  uint64_t src_mask;
  uint32_t scale;          // output multiplier (1 = none)
  uint64_t dst_stride, src_stride;  // words between the columns of a batch (blockIdx.y)
  uint32_t* const* cols;   // NTT_COLS: device array of column pointers (destinations)
  const uint32_t* const* src_cols;  // NTT_COLS: source columns, or null to read the destinations
};

// stwo circle twiddle of circle butterfly h from line layer 0.
__device__ __forceinline__ uint32_t stwo_circle_tw(const uint32_t* l0, uint64_t h) {
  const uint32_t x = l0[(h >> 1) ^ 1];
  return ((h ^ (h >> 1)) & 1) ? m31_neg(x) : x;
}

template <bool FWD, int MODE>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_pass(uint32_t* dst, const uint32_t* src, PassArgs a) {
  extern __shared__ uint32_t lds[];
  if constexpr ((MODE & NTT_COLS) != 0) {
    dst = ld_uniform(a.cols + blockIdx.y);
    src = a.src_cols ? ld_uniform(a.src_cols + blockIdx.y) : dst;
  } else {
    dst += blockIdx.y * a.dst_stride;
    src += blockIdx.y * a.src_stride;
  }
  const uint32_t cmask = (1u << a.cbits) - 1;
  const uint32_t E = 1u << (a.k + a.cbits);
  const uint32_t low_tiles_log = a.j0 - a.cbits;
  const uint64_t tile = blockIdx.x;
  const uint64_t low_base = (tile & ((1ull << low_tiles_log) - 1)) << a.cbits;
  const uint64_t hi = tile >> low_tiles_log;
  const uint64_t hi_base = hi << (a.j0 + a.k);

  for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
    uint64_t gi = hi_base | ((uint64_t)(e >> a.cbits) << a.j0) | (low_base + (e & cmask));
    lds[e] = src[gi & a.src_mask];
  }
  __syncthreads();

  const uint32_t half = E >> 1;
  for (uint32_t s = 0; s < a.k; s++) {
    const uint32_t b = FWD ? (a.k - 1 - s) : s;
    const uint32_t* tw = a.tw[b];
    const uint32_t bmask = (1u << b) - 1;
    for (uint32_t u = threadIdx.x; u < half; u += blockDim.x) {
      const uint32_t c = u & cmask, rr = u >> a.cbits;
      const uint32_t r0 = ((rr >> b) << (b + 1)) | (rr & bmask);
      const uint64_t h = (hi << (a.k - b - 1)) | (r0 >> (b + 1));
      uint32_t w;
      if constexpr ((MODE & NTT_STWO) != 0)
        w = a.j0 + b == 0 ? stwo_circle_tw(tw, h) : tw[h];
      else
        w = tw[h];
      const uint32_t i0 = (r0 << a.cbits) | c, i1 = ((r0 | (1u << b)) << a.cbits) | c;
      const uint32_t v0 = lds[i0], v1 = lds[i1];
      if (FWD) {
        const uint32_t t = m31_mul(v1, w);
        lds[i0] = m31_add(v0, t);
        lds[i1] = m31_sub(v0, t);
      } else {
        lds[i0] = m31_add(v0, v1);
        lds[i1] = m31_mul(m31_sub(v0, v1), w);
      }
    }
    __syncthreads();
  }

  for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
    uint64_t gi = hi_base | ((uint64_t)(e >> a.cbits) << a.j0) | (low_base + (e & cmask));
    uint32_t v = lds[e];
    if (a.scale != 1u) v = m31_mul(v, a.scale);
    dst[gi] = v;
  }
}

// ---------------------------------------------------------------------------
// Register-radix pass for 16384-element tiles (k + cbits = 14): 512 lanes x
// 32 elements. A round butterflies up to 5 bits in VGPRs: lane t holds the 32
// elements whose tile index e has bits [pos, pos+5) = s (register index) and
// the other 9 bits = t. Between rounds the tile moves through LDS once, with
// an XOR swizzle that keeps all three access patterns bank-conflict free.
// For register bit q of a round the twiddle index is
//   h = (H << (13 - pos - q)) | ((t >> pos) << (4 - q)) | (s0 >> (q + 1)),
// so a lane needs only the 2^(4-q) consecutive words starting at the first
// term; they are loaded once per round instead of once per butterfly.
static constexpr uint32_t RX_THREADS = 512;
static constexpr uint32_t RX_E_BITS = 14;

struct RadixRound {
  uint32_t rb0, nb, pos;  // row bits [rb0, rb0 + nb) at register window [pos, pos + 5)
};
struct RadixArgs {
  const uint32_t* tw[16];  // twiddle table of row bit b (global bit j0 + b)
  uint32_t j0, k, cbits;
  uint32_t n_rounds;
  RadixRound round[4];     // in execution order
  uint64_t src_mask;
  uint32_t scale;
  uint64_t dst_stride, src_stride;
  uint32_t* const* cols;   // NTT_COLS: device array of column pointers (destinations)
  const uint32_t* const* src_cols;  // NTT_COLS: source columns, or null to read the destinations
  uint32_t launch_order;   // 1: tile b runs as workgroup b (else XCD-contiguous ranges)
};

__device__ __forceinline__ uint32_t rx_own(uint32_t t, uint32_t s, uint32_t pos) {
  return (t & ((1u << pos) - 1)) | (s << pos) | ((t >> pos) << (pos + 5));
}
__device__ __forceinline__ uint32_t rx_swz(uint32_t e) { return e ^ ((e >> 5) & 31u); }

// Twiddles of register bit Q for this lane: the 2^(4-Q) consecutive words
// from `base` (a multiple of 2^(4-Q)); 16-byte loads when aligned.
template <int Q>
__device__ __forceinline__ void rx_load_tw(const uint32_t* table, uint64_t base, uint32_t (&w)[1u << (4 - Q)]) {
  constexpr uint32_t NW = 1u << (4 - Q);
  const uint32_t* p = table + base;
  if constexpr (NW >= 4) {
    if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
#pragma unroll
      for (uint32_t m = 0; m < NW; m += 4) {
        const uint4 x = *reinterpret_cast<const uint4*>(p + m);
        w[m] = x.x;
        w[m + 1] = x.y;
        w[m + 2] = x.z;
        w[m + 3] = x.w;
      }
      return;
    }
  }
#pragma unroll
  for (uint32_t m = 0; m < NW; m++) w[m] = p[m];
}

// stwo circle bit (register bit 0 at tile bit 0): the 16 circle twiddles
// from `base` (a multiple of 16) are +-L0[base / 2 + ((m >> 1) ^ 1)], m < 16,
// negated when m & 3 is 1 or 2: 8 consecutive words of line layer 0.
__device__ __forceinline__ void rx_load_stwo_circle(const uint32_t* l0, uint64_t base, uint32_t (&w)[16]) {
  uint32_t x[8];
  rx_load_tw<1>(l0, base >> 1, x);
#pragma unroll
  for (uint32_t m = 0; m < 16; m++) {
    const uint32_t v = x[(m >> 1) ^ 1];
    w[m] = ((m ^ (m >> 1)) & 1) ? m31_neg(v) : v;
  }
}

template <int Q, int MODE>
__device__ __forceinline__ void rx_load_bit(const uint32_t* table, uint32_t epos, uint64_t base,
                                            uint32_t (&w)[1u << (4 - Q)]) {
  if constexpr (Q == 0 && (MODE & NTT_STWO) != 0) {
    if (epos == 0) {
      rx_load_stwo_circle(table, base, w);
      return;
    }
  }
  rx_load_tw<Q>(table, base, w);
}

template <bool FWD, int Q>
__device__ __forceinline__ void rx_apply(uint32_t (&v)[32], const uint32_t (&w)[1u << (4 - Q)]) {
#pragma unroll
  for (uint32_t s0 = 0; s0 < 32; s0++) {
    if (s0 & (1u << Q)) continue;
    const uint32_t s1 = s0 | (1u << Q);
    const uint32_t tw = w[s0 >> (Q + 1)];
    if (FWD) {
      const uint32_t m = m31_mul(v[s1], tw);
      v[s1] = m31_sub(v[s0], m);
      v[s0] = m31_add(v[s0], m);
    } else {
      const uint32_t d = m31_sub(v[s0], v[s1]);
      v[s0] = m31_add(v[s0], v[s1]);
      v[s1] = m31_mul(d, tw);
    }
  }
}

// Word at a 32-bit byte offset from a wave-uniform base: the compiler
// addresses it as scalar base + one VGPR offset, so an element costs one
// 32-bit add instead of 64-bit index arithmetic.
__device__ __forceinline__ uint32_t ld_at(const uint32_t* base, uint32_t byte_off) {
  return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void st_at(uint32_t* base, uint32_t byte_off, uint32_t v) {
  *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// Pass arguments with the round layout fixed at compile time (defined
// below); their tiles load a round's twiddles before its butterflies.
template <bool FWD, int SHAPE>
struct FixedArgs;
template <typename A>
struct rx_is_fixed {
  static constexpr bool value = false;
};
template <bool F, int S>
struct rx_is_fixed<FixedArgs<F, S>> {
  static constexpr bool value = true;
};
// Fixed strided shapes (k row bits x 2^(14-k) columns, shape id 16 + k)
// address their global elements from a uniform tile base (ld_at / st_at).
static constexpr int RX_STRIDED_K0 = 16;
template <typename A>
struct rx_is_strided {
  static constexpr bool value = false;
};
template <bool F, int S>
struct rx_is_strided<FixedArgs<F, S>> {
  static constexpr bool value = S > RX_STRIDED_K0;
};

// Word offset of tile element e in HBM: column bits [0, cbits) stay, row
// bits move up to [j0, j0 + k). The map is linear over disjoint bit ranges,
// so off(e0 | (s << pos)) = off(e0) + off(s << pos).
__device__ __forceinline__ uint32_t rx_goff(uint32_t e, uint32_t cbits, uint32_t j0) {
  return ((e >> cbits) << j0) | (e & ((1u << cbits) - 1));
}

// One 16384-element tile of a pass; `a` may live in kernel arguments or in
// global memory (work-list jobs), `lds` holds 2^14 words.
template <bool FWD, int MODE, typename Args>
__device__ __forceinline__ void radix_tile(uint32_t* dst, const uint32_t* src, const Args& a, uint64_t tile,
                                           uint32_t* lds) {
  const uint32_t t = threadIdx.x;
  const uint32_t cmask = (1u << a.cbits) - 1;
  const uint32_t low_tiles_log = a.j0 - a.cbits;
  const uint64_t low_base = (tile & ((1ull << low_tiles_log) - 1)) << a.cbits;
  const uint64_t H = tile >> low_tiles_log;
  const uint64_t hi_base = H << (a.j0 + a.k);
  constexpr bool kStrided = rx_is_strided<Args>::value;
  // One round of the pass; the fixed shapes unroll their constant round
  // count, the generic kernel loops over its runtime one.
  auto round = [&](uint32_t ri) __attribute__((always_inline)) {
    uint32_t v[32];
    const uint32_t pos = a.round[ri].pos, r_rb0 = a.round[ri].rb0, r_nb = a.round[ri].nb;
    // A contiguous pass (cbits = 0) whose round keeps each lane's 32 elements
    // next to each other (pos < 4) would read or write HBM with a 128-byte
    // stride between lanes; it goes through LDS instead, so that every
    // 16-byte access of a wavefront is part of one contiguous 1 KiB run.
    const bool staged = a.cbits == 0 && a.j0 == 0 && pos < 4;
    // LDS word of element s: swz(own(t, 0, pos)) ^ swz(s << pos) (own's bit
    // groups are disjoint, so the swizzle splits over them); the second term
    // is wave-uniform, so each access costs one XOR with a scalar.
    const uint32_t lb = rx_swz(rx_own(t, 0, pos));
    if (ri == 0 && staged) {
#pragma unroll
      for (uint32_t g = 0; g < 8; g++) {
        const uint32_t e = (g << 11) | (t << 2);
        const uint4 x = *reinterpret_cast<const uint4*>(src + ((hi_base | e) & a.src_mask));
        lds[rx_swz(e)] = x.x;
        lds[rx_swz(e + 1)] = x.y;
        lds[rx_swz(e + 2)] = x.z;
        lds[rx_swz(e + 3)] = x.w;
      }
      __syncthreads();
#pragma unroll
      for (uint32_t s = 0; s < 32; s++) v[s] = lds[lb ^ rx_swz(s << pos)];
    } else if (ri == 0 && kStrided) {
      // element s sits off(s << pos) words after the lane's first one (a
      // wave-uniform step); the tile spans < 2^32 bytes and masking the
      // source index only clears bits of hi_base (both checked when the
      // shape is chosen)
      const uint32_t* sp = src + ((hi_base & a.src_mask) | low_base);
      const uint32_t lane = rx_goff(rx_own(t, 0, pos), a.cbits, a.j0) << 2;
#pragma unroll
      for (uint32_t s = 0; s < 32; s++) v[s] = ld_at(sp, lane + (rx_goff(s << pos, a.cbits, a.j0) << 2));
    } else if (ri == 0) {
#pragma unroll
      for (uint32_t s = 0; s < 32; s++) {
        const uint32_t e = rx_own(t, s, pos);
        const uint64_t gi = hi_base | ((uint64_t)(e >> a.cbits) << a.j0) | (low_base + (e & cmask));
        v[s] = src[gi & a.src_mask];
      }
    } else {
#pragma unroll
      for (uint32_t s = 0; s < 32; s++) v[s] = lds[lb ^ rx_swz(s << pos)];
    }
    const uint32_t qlo = r_rb0 + a.cbits - pos, qhi = qlo + r_nb;  // register bits to butterfly
    // The generic kernel loads each bit's twiddles just before its
    // butterflies. (Loading the whole round's 31 words up front overlaps
    // their latencies within a workgroup but costs 26 more VGPRs: 154 per
    // lane fits one 512-lane workgroup per CU, 128 fits the two that LDS
    // allows, and the second workgroup hides the latency better: the LDE ran
    // 0.60 -> 0.41 ms.) The fixed-shape kernels need about 100 VGPRs, so
    // they load the round's twiddles up front and still fit two workgroups.
    constexpr bool PRELOAD = rx_is_fixed<Args>::value;
    uint32_t w4[1], w3[2], w2[4], w1[8], w0[16];
#define VX_RX_LOAD(Q, W)                                                                          \
  if ((Q) >= qlo && (Q) < qhi) {                                                                  \
    const uint32_t epos = pos + (Q);                                                              \
    const uint64_t base = (H << (13 - epos)) | ((uint64_t)(t >> pos) << (4 - (Q)));               \
    rx_load_bit<(Q), MODE>(a.tw[epos - a.cbits], epos, base, W);                                  \
  }
#define VX_RX_APPLY(Q, W) \
  if ((Q) >= qlo && (Q) < qhi) rx_apply<FWD, (Q)>(v, W);
    if (PRELOAD && FWD) {
      VX_RX_LOAD(4, w4) VX_RX_LOAD(3, w3) VX_RX_LOAD(2, w2) VX_RX_LOAD(1, w1) VX_RX_LOAD(0, w0)
      VX_RX_APPLY(4, w4) VX_RX_APPLY(3, w3) VX_RX_APPLY(2, w2) VX_RX_APPLY(1, w1) VX_RX_APPLY(0, w0)
    } else if (PRELOAD) {
      VX_RX_LOAD(0, w0) VX_RX_LOAD(1, w1) VX_RX_LOAD(2, w2) VX_RX_LOAD(3, w3) VX_RX_LOAD(4, w4)
      VX_RX_APPLY(0, w0) VX_RX_APPLY(1, w1) VX_RX_APPLY(2, w2) VX_RX_APPLY(3, w3) VX_RX_APPLY(4, w4)
    } else if (FWD) {
      VX_RX_LOAD(4, w4) VX_RX_APPLY(4, w4) VX_RX_LOAD(3, w3) VX_RX_APPLY(3, w3) VX_RX_LOAD(2, w2)
      VX_RX_APPLY(2, w2) VX_RX_LOAD(1, w1) VX_RX_APPLY(1, w1) VX_RX_LOAD(0, w0) VX_RX_APPLY(0, w0)
    } else {
      VX_RX_LOAD(0, w0) VX_RX_APPLY(0, w0) VX_RX_LOAD(1, w1) VX_RX_APPLY(1, w1) VX_RX_LOAD(2, w2)
      VX_RX_APPLY(2, w2) VX_RX_LOAD(3, w3) VX_RX_APPLY(3, w3) VX_RX_LOAD(4, w4) VX_RX_APPLY(4, w4)
    }
#undef VX_RX_APPLY
#undef VX_RX_LOAD
    if (ri + 1 < a.n_rounds) {
#pragma unroll
      for (uint32_t s = 0; s < 32; s++) lds[lb ^ rx_swz(s << pos)] = v[s];
      __syncthreads();
    } else if (staged) {
      // each lane rewrites exactly the words it read this round, so no
      // barrier is needed before these stores
      if (!FWD && a.scale != 1u) {
#pragma unroll
        for (uint32_t s = 0; s < 32; s++) lds[lb ^ rx_swz(s << pos)] = m31_mul(v[s], a.scale);
      } else {
#pragma unroll
        for (uint32_t s = 0; s < 32; s++) lds[lb ^ rx_swz(s << pos)] = v[s];
      }
      __syncthreads();
#pragma unroll
      for (uint32_t g = 0; g < 8; g++) {
        const uint32_t e = (g << 11) | (t << 2);
        *reinterpret_cast<uint4*>(dst + (hi_base | e)) =
            make_uint4(lds[rx_swz(e)], lds[rx_swz(e + 1)], lds[rx_swz(e + 2)], lds[rx_swz(e + 3)]);
      }
    } else if (kStrided) {
      uint32_t* dp = dst + (hi_base | low_base);
      const uint32_t lane = rx_goff(rx_own(t, 0, pos), a.cbits, a.j0) << 2;
      if (!FWD && a.scale != 1u) {
#pragma unroll
        for (uint32_t s = 0; s < 32; s++)
          st_at(dp, lane + (rx_goff(s << pos, a.cbits, a.j0) << 2), m31_mul(v[s], a.scale));
      } else {
#pragma unroll
        for (uint32_t s = 0; s < 32; s++) st_at(dp, lane + (rx_goff(s << pos, a.cbits, a.j0) << 2), v[s]);
      }
    } else {
#pragma unroll
      for (uint32_t s = 0; s < 32; s++) {
        const uint32_t e = rx_own(t, s, pos);
        const uint64_t gi = hi_base | ((uint64_t)(e >> a.cbits) << a.j0) | (low_base + (e & cmask));
        dst[gi] = (!FWD && a.scale != 1u) ? m31_mul(v[s], a.scale) : v[s];
      }
    }
  };
  if constexpr (rx_is_fixed<Args>::value) {
#pragma unroll
    for (uint32_t ri = 0; ri < Args::n_rounds; ri++) round(ri);
  } else {
    for (uint32_t ri = 0; ri < a.n_rounds; ri++) round(ri);
  }
}

// Workgroup b runs on XCD b mod 8 (round-robin dispatch). A strided pass's
// tiles are 64-byte row segments, so tiles t and t ^ 1 share every 128-byte
// line; giving each XCD a contiguous range of tiles makes the two run back
// to back on one XCD, where the second finds the lines in that XCD's L2.
__device__ __forceinline__ uint64_t xcd_tile(uint32_t b, uint32_t grid) {
  return (grid & 7u) == 0 ? (uint64_t)(b & 7u) * (grid >> 3) + (b >> 3) : b;
}

// The two pass shapes every large transform uses, with their rounds as
// compile-time constants (shifts, swizzles and round structure fold away;
// the contiguous pass also drops from 128 to about 100 VGPRs):
//   RX_CONTIG   bits 0..13 (j0 = 0, cbits = 0): rounds (rb0, nb, pos) =
//               (9, 5, 9), (4, 5, 4), (0, 4, 0) for evaluate, reversed for
//               interpolate (make_radix_args' schedule);
//   RX_STRIDED  k row bits x 2^(14-k) columns (k = 1..10, j0 from the
//               pass; the kernel switches on k): for k = 10, (5, 5, 9),
//               (0, 5, 4), reversed for interpolate.
// A/B against the generic kernel on one box: 2^24 stwo evaluate 0.125 ->
// 0.111 ms, Fibonacci LDE (2^24 -> 2^26) 0.412 -> 0.346 ms, the batched
// 96-column LDE 2.23 -> 2.02 ms. The work-list kernel uses the contiguous
// shape for its first and last stages (every column's 14-bit pass). Loading
// each round's twiddles up front (radix_tile) then took the Fibonacci LDE
// 0.350 -> 0.338 ms.
enum : int { RX_GENERIC = 0, RX_CONTIG = 1, RX_STRIDED = 2 };

template <bool FWD>
struct FixedArgs<FWD, RX_CONTIG> {
  static constexpr uint32_t k = 14, cbits = 0, j0 = 0, n_rounds = 3;
  static constexpr RadixRound round[3] = {FWD ? RadixRound{9, 5, 9} : RadixRound{0, 4, 0}, RadixRound{4, 5, 4},
                                          FWD ? RadixRound{0, 4, 0} : RadixRound{9, 5, 9}};
  const uint32_t* tw[16];
  uint64_t src_mask;
  uint32_t scale;
  template <typename A>
  __device__ explicit FixedArgs(const A& a) : src_mask(a.src_mask), scale(a.scale) {
#pragma unroll
    for (int i = 0; i < 16; i++) tw[i] = a.tw[i];
  }
};
// Strided shape with K row bits (1 <= K <= 10) and 14 - K columns, j0 from
// the pass; make_radix_args' schedule with constant rounds: K <= 5 is one
// round (0, K, 9); K > 5 is (K - 5, 5, 9) then (0, K - 5, 14 - K) for
// evaluate, reversed for interpolate (K = 10 is the 16-column shape above).
template <bool FWD, int SHAPE>
struct FixedArgs {
  static constexpr uint32_t k = SHAPE - RX_STRIDED_K0, cbits = RX_E_BITS - k, n_rounds = k > 5 ? 2 : 1;
  static_assert(k >= 1 && k <= 10, "strided shapes have 1..10 row bits");
  static constexpr RadixRound hi_r = k > 5 ? RadixRound{k - 5, 5, 9} : RadixRound{0, k, 9};
  static constexpr RadixRound lo_r = RadixRound{0, k > 5 ? k - 5 : 0, cbits};
  static constexpr RadixRound round[2] = {n_rounds == 1 || FWD ? hi_r : lo_r, FWD ? lo_r : hi_r};
  uint32_t j0;
  const uint32_t* tw[16];
  uint64_t src_mask;
  uint32_t scale;
  template <typename A>
  __device__ explicit FixedArgs(const A& a) : j0(a.j0), src_mask(a.src_mask), scale(a.scale) {
#pragma unroll
    for (int i = 0; i < 16; i++) tw[i] = a.tw[i];
  }
};

// A strided tile of any row count: the pass's k is uniform over the launch
// (or over the work-list job), so the switch costs a few scalar instructions.
template <bool FWD, int MODE, typename A>
__device__ __forceinline__ void radix_tile_strided(uint32_t* dst, const uint32_t* src, const A& a, uint64_t tile,
                                                   uint32_t* lds) {
#define VX_RX_K(K)                                                  \
  case K: {                                                         \
    const FixedArgs<FWD, RX_STRIDED_K0 + K> f(a);                   \
    radix_tile<FWD, MODE>(dst, src, f, tile, lds);                  \
    break;                                                          \
  }
  switch (a.k) {
    VX_RX_K(1) VX_RX_K(2) VX_RX_K(3) VX_RX_K(4) VX_RX_K(5)
    VX_RX_K(6) VX_RX_K(7) VX_RX_K(8) VX_RX_K(9) VX_RX_K(10)
    default: break;
  }
#undef VX_RX_K
}

// Strided passes (cbits = 14 - k columns) take the fixed shapes when their
// 32-bit element offsets fit (tile span below 2^32 bytes) and the source
// mask keeps every in-tile bit, so masking touches only the tile base.
static bool rx_strided_ok(uint32_t j0, uint32_t k, uint32_t cbits, uint64_t src_mask) {
  return k >= 1 && k <= 10 && cbits == RX_E_BITS - k && cbits <= j0 && j0 + k <= 30 &&
         src_mask >= (1ull << (j0 + k)) - 1;
}

static int radix_shape(const RadixArgs& r) {
  if (r.k == 14 && r.cbits == 0 && r.j0 == 0) return RX_CONTIG;
  if (rx_strided_ok(r.j0, r.k, r.cbits, r.src_mask)) return RX_STRIDED;
  return RX_GENERIC;
}

template <bool FWD, int MODE, int SHAPE>
__global__ __launch_bounds__(RX_THREADS) void k_ntt_radix(uint32_t* dst, const uint32_t* src, RadixArgs a) {
  __shared__ uint32_t lds[1u << RX_E_BITS];
  if constexpr ((MODE & NTT_COLS) != 0) {
    dst = ld_uniform(a.cols + blockIdx.y);
    src = a.src_cols ? ld_uniform(a.src_cols + blockIdx.y) : dst;
  } else {
    dst += blockIdx.y * a.dst_stride;
    src += blockIdx.y * a.src_stride;
  }
  const uint64_t tile = a.launch_order ? (uint64_t)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
  if constexpr (SHAPE == RX_GENERIC) {
    radix_tile<FWD, MODE>(dst, src, a, tile, lds);
  } else if constexpr (SHAPE == RX_STRIDED) {
    radix_tile_strided<FWD, MODE>(dst, src, a, tile, lds);
  } else {
    const FixedArgs<FWD, SHAPE> f(a);
    radix_tile<FWD, MODE>(dst, src, f, tile, lds);
  }
}

// Work list: the tiles of many passes (columns of different sizes) in one
// launch. Workgroup b runs tile b - first_tile of the last job whose
// first_tile <= b (jobs sorted by first_tile).
struct RadixJob {
  uint32_t* dst;
  const uint32_t* src;
  uint64_t first_tile;
  RadixArgs a;
};

template <bool FWD, int SHAPE>
__global__ __launch_bounds__(RX_THREADS) void k_ntt_radix_list(const RadixJob* jobs, uint32_t n_jobs) {
  __shared__ uint32_t lds[1u << RX_E_BITS];
  const uint64_t b = blockIdx.x;
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ld_uniform(&jobs[mid].first_tile) <= b) lo = mid;
    else hi = mid;
  }
  uniform_t<RadixJob>& J = as_uniform(jobs + lo);
  if constexpr (SHAPE == RX_GENERIC) {
    radix_tile<FWD, 0>(J.dst, J.src, J.a, b - J.first_tile, lds);
  } else if constexpr (SHAPE == RX_STRIDED) {
    radix_tile_strided<FWD, 0>(J.dst, J.src, J.a, b - J.first_tile, lds);
  } else {
    const FixedArgs<FWD, SHAPE> f(J.a);
    radix_tile<FWD, 0>(J.dst, J.src, f, b - J.first_tile, lds);
  }
}

static const uint32_t* bit_table(const TwiddleRef& tw, uint32_t j) {
  if (j == 0) return tw.stwo_circle ? tw.line + tw.off[0] : tw.circle;
  return tw.line + tw.off[j - 1];
}

struct Batch {
  uint32_t n_cols;
  uint64_t dst_stride, src_stride;
  uint32_t* const* cols;  // device column pointers, or null for strided columns
  const uint32_t* const* src_cols;  // with cols: the first pass's source columns (null: in place)
};

using RadixKernel = void (*)(uint32_t*, const uint32_t*, RadixArgs);
using PassKernel = void (*)(uint32_t*, const uint32_t*, PassArgs);
#define VX_RX_MODES(F, S) {k_ntt_radix<F, 0, S>, k_ntt_radix<F, 1, S>, k_ntt_radix<F, 2, S>, k_ntt_radix<F, 3, S>}
static RadixKernel radix_kernel(bool fwd, int mode, int shape) {
  static const RadixKernel k[2][3][4] = {
      {VX_RX_MODES(false, RX_GENERIC), VX_RX_MODES(false, RX_CONTIG), VX_RX_MODES(false, RX_STRIDED)},
      {VX_RX_MODES(true, RX_GENERIC), VX_RX_MODES(true, RX_CONTIG), VX_RX_MODES(true, RX_STRIDED)}};
  return k[fwd ? 1 : 0][shape][mode];
}
#undef VX_RX_MODES
static PassKernel pass_kernel(bool fwd, int mode) {
  static const PassKernel k[2][4] = {
      {k_ntt_pass<false, 0>, k_ntt_pass<false, 1>, k_ntt_pass<false, 2>, k_ntt_pass<false, 3>},
      {k_ntt_pass<true, 0>, k_ntt_pass<true, 1>, k_ntt_pass<true, 2>, k_ntt_pass<true, 3>}};
  return k[fwd ? 1 : 0][mode];
}

// Arguments of a 16384-element radix pass over bits [j0, j0 + k) with
// 2^cbits columns per row: rounds of <= 5 row bits from the top; evaluate
// runs them top-down, interpolate bottom-up.
static RadixArgs make_radix_args(bool fwd, const TwiddleRef& tw, uint32_t j0, uint32_t k, uint32_t cbits,
                                 uint64_t src_mask, uint32_t scale) {
  require(k >= 1 && k <= 14 && k + cbits == RX_E_BITS && cbits <= j0 && j0 + k <= tw.log_n,
          "ntt: invalid radix pass shape");
  RadixArgs r{};
  for (uint32_t b = 0; b < k; b++) r.tw[b] = bit_table(tw, j0 + b);
  r.j0 = j0;
  r.k = k;
  r.cbits = cbits;
  r.src_mask = src_mask;
  r.scale = scale;
  RadixRound rr[4];
  uint32_t nr = 0;
  for (int32_t hi = (int32_t)k; hi > 0; hi -= 5) {
    uint32_t rb0 = hi >= 5 ? (uint32_t)hi - 5 : 0;
    uint32_t nb = (uint32_t)hi - rb0;
    uint32_t pos = rb0 + cbits < 9 ? rb0 + cbits : 9;
    rr[nr++] = {rb0, nb, pos};
  }
  r.n_rounds = nr;
  for (uint32_t i = 0; i < nr; i++) r.round[i] = fwd ? rr[i] : rr[nr - 1 - i];
  return r;
}

static void launch_pass(bool fwd, uint32_t* dst, const uint32_t* src, uint64_t src_mask,
                        const TwiddleRef& tw, uint32_t j0, uint32_t k, uint32_t cbits, uint32_t scale,
                        const Batch& bt, hipStream_t s, uint32_t launch_order = 0) {
  const uint32_t log_n = tw.log_n;
  require(k >= 1 && k <= 16 && k + cbits <= NTT_TILE_LOG && k + cbits <= log_n && cbits <= j0 &&
              j0 + k <= log_n,
          "ntt: invalid pass shape");
  PassArgs a{};
  for (uint32_t b = 0; b < k; b++) a.tw[b] = bit_table(tw, j0 + b);
  a.j0 = j0;
  a.k = k;
  a.cbits = cbits;
  a.src_mask = src_mask;
  a.scale = scale;
  a.dst_stride = bt.dst_stride;
  a.src_stride = bt.src_stride;
  a.cols = bt.cols;
  a.src_cols = bt.src_cols;
  require(bt.n_cols >= 1 && bt.n_cols <= 65535, "ntt: 1..65535 columns per batch");
  const int mode = (tw.stwo_circle ? NTT_STWO : 0) | (bt.cols ? NTT_COLS : 0);
  const uint32_t e_log = k + cbits;
  const uint64_t grid = 1ull << (log_n - e_log);
  if (e_log == RX_E_BITS) {
    RadixArgs r = make_radix_args(fwd, tw, j0, k, cbits, src_mask, scale);
    r.dst_stride = bt.dst_stride;
    r.src_stride = bt.src_stride;
    r.cols = bt.cols;
    r.src_cols = bt.src_cols;
    r.launch_order = launch_order;
    hipLaunchKernelGGL(radix_kernel(fwd, mode, radix_shape(r)), dim3((uint32_t)grid, bt.n_cols), dim3(RX_THREADS), 0,
                       s, dst, src, r);
    VX_HIP(hipGetLastError());
    return;
  }
  const uint32_t threads = e_log >= 10 ? NTT_THREADS : (e_log >= 7 ? (1u << (e_log - 1)) : 64u);
  const size_t lds = (size_t)4 << e_log;
  hipLaunchKernelGGL(pass_kernel(fwd, mode), dim3((uint32_t)grid, bt.n_cols), dim3(threads), lds, s, dst, src, a);
  VX_HIP(hipGetLastError());
}

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Strided passes cover [NTT_TILE_LOG, eff_bits) from the top, then one
// contiguous pass covers the low bits. The plan (tuned per transform shape,
// tune.cpp) may split the strided bits over one more pass (fewer rows, wider
// column runs per tile) and may run tiles in launch order instead of
// XCD-contiguous ranges; every plan computes the same transform.
static uint32_t strided_passes(uint32_t strided_bits, const NttPlan& plan) {
  uint32_t n = (strided_bits + NTT_STRIDED_MAX - 1) / NTT_STRIDED_MAX;
  if (plan.split && n && strided_bits > n) n++;
  return n;
}

static void forward_passes(uint32_t* data, const uint32_t* src, uint64_t src_mask, const TwiddleRef& tw,
                           uint32_t eff_bits, hipStream_t s, Batch bt, const NttPlan& plan) {
  const uint32_t log_n = tw.log_n;
  const uint64_t dst_stride = bt.dst_stride;
  uint32_t top = eff_bits;
  const uint32_t contig = eff_bits < NTT_TILE_LOG ? eff_bits : NTT_TILE_LOG;
  const uint32_t strided_bits = eff_bits - contig;
  const uint32_t n_strided = strided_passes(strided_bits, plan);
  const uint32_t* cur = src;
  uint64_t mask = src_mask;
  for (uint32_t p = 0; p < n_strided; p++) {
    uint32_t k = (top - contig) / (n_strided - p);  // balanced chunks
    uint32_t j0 = top - k;
    launch_pass(true, data, cur, mask, tw, j0, k, NTT_TILE_LOG - k, 1u, bt, s, plan.order);
    top = j0;
    cur = data;
    mask = (1ull << log_n) - 1;
    bt.src_stride = dst_stride;
    bt.src_cols = nullptr;  // later passes work in place on the destinations
  }
  launch_pass(true, data, cur, mask, tw, 0, contig, 0, 1u, bt, s, plan.order);
}

void ntt_forward(uint32_t* data, const uint32_t* src, uint64_t src_mask, const TwiddleRef& tw,
                 uint32_t eff_bits, hipStream_t s, uint32_t n_cols, uint64_t dst_stride, uint64_t src_stride) {
  const uint32_t log_n = tw.log_n;
  require(aligned16(data) && aligned16(src) && dst_stride % 4 == 0 && src_stride % 4 == 0,
          "ntt: columns must be 16-byte aligned");
  require(log_n >= 1 && log_n <= 31 && eff_bits >= 1 && eff_bits <= log_n, "ntt_forward: bad sizes");
  require(((src_mask + 1) & src_mask) == 0 && src_mask + 1 >= (1ull << eff_bits),
          "ntt_forward: source length must be a power of two covering the transformed bits");
  require(n_cols == 1 || (dst_stride >= (1ull << log_n) && src_stride >= src_mask + 1),
          "ntt_forward: batch columns must not overlap");
  const Batch bt{n_cols, dst_stride, src_stride, nullptr, nullptr};
  const NttKey key{true, tw.log_n, eff_bits, n_cols, tw.stwo_circle != 0};
  // out of place, the first pass reads src intact, so the candidates run on
  // the caller's buffers; in place they run on a scratch buffer of the same
  // shape (allocated only when this shape is measured)
  struct Scratch {
    void* p = nullptr;
    ~Scratch() {
      if (p) (void)hipFree(p);
    }
  } scratch;
  const uint64_t words = (n_cols > 1 ? dst_stride : 0) * (n_cols - 1) + (1ull << tw.log_n);
  const NttPlan plan = tuned_ntt_plan(key, s, [&](const NttPlan& p) {
    if (src != data) {
      forward_passes(data, src, src_mask, tw, eff_bits, s, bt, p);
      return true;
    }
    if (words > NTT_TUNE_SCRATCH_MAX_WORDS) return false;
    if (!scratch.p) VX_HIP(hipMalloc(&scratch.p, 4 * words));
    uint32_t* t = static_cast<uint32_t*>(scratch.p);
    forward_passes(t, t, src_mask, tw, eff_bits, s, bt, p);
    return true;
  });
  forward_passes(data, src, src_mask, tw, eff_bits, s, bt, plan);
}

void ntt_forward_cols(uint32_t* const* d_cols, uint32_t n_cols, const TwiddleRef& tw, hipStream_t s,
                      const uint32_t* const* d_src_cols, uint32_t src_log_n) {
  require(d_cols != nullptr && tw.log_n >= 1 && tw.log_n <= 31, "ntt_forward_cols: bad arguments");
  const uint32_t eff = d_src_cols ? src_log_n : tw.log_n;
  require(eff >= 1 && eff <= tw.log_n, "ntt_forward_cols: source size must be 2..2^log_n");
  const NttPlan plan = known_ntt_plan(NttKey{true, tw.log_n, eff, n_cols, tw.stwo_circle != 0});
  forward_passes(nullptr, nullptr, (1ull << eff) - 1, tw, eff, s, Batch{n_cols, 0, 0, d_cols, d_src_cols}, plan);
}

static void inverse_passes(uint32_t* data, const TwiddleRef& tw, hipStream_t s, const Batch& bt,
                           const NttPlan& plan) {
  const uint32_t log_n = tw.log_n;
  const uint32_t e = (30u * log_n) % 31u;
  const uint32_t inv_n = e == 0 ? 1u : (1u << e);
  const uint64_t mask = (1ull << log_n) - 1;
  const uint32_t contig = log_n < NTT_TILE_LOG ? log_n : NTT_TILE_LOG;
  const uint32_t strided_bits = log_n - contig;
  const uint32_t n_strided = strided_passes(strided_bits, plan);
  launch_pass(false, data, data, mask, tw, 0, contig, 0, n_strided == 0 ? inv_n : 1u, bt, s, plan.order);
  uint32_t bot = contig;
  for (uint32_t p = 0; p < n_strided; p++) {
    uint32_t k = (log_n - bot) / (n_strided - p);
    bool last = p + 1 == n_strided;
    launch_pass(false, data, data, mask, tw, bot, k, NTT_TILE_LOG - k, last ? inv_n : 1u, bt, s, plan.order);
    bot += k;
  }
}

void ntt_inverse(uint32_t* data, const TwiddleRef& tw, hipStream_t s, uint32_t n_cols, uint64_t stride) {
  const uint32_t log_n = tw.log_n;
  require(aligned16(data) && stride % 4 == 0, "ntt: columns must be 16-byte aligned");
  require(log_n >= 1 && log_n <= 31, "ntt_inverse: bad size");
  require(n_cols == 1 || stride >= (1ull << log_n), "ntt_inverse: batch columns must not overlap");
  const Batch bt{n_cols, stride, stride, nullptr, nullptr};
  const NttKey key{false, log_n, log_n, n_cols, tw.stwo_circle != 0};
  // in place: the candidates are timed on a scratch buffer of the same
  // shape (allocated only when this shape is measured)
  struct Scratch {
    void* p = nullptr;
    ~Scratch() {
      if (p) (void)hipFree(p);
    }
  } scratch;
  const uint64_t words = (n_cols > 1 ? stride : 0) * (n_cols - 1) + (1ull << log_n);
  const NttPlan plan = tuned_ntt_plan(key, s, [&](const NttPlan& p) {
    if (words > NTT_TUNE_SCRATCH_MAX_WORDS) return false;
    if (!scratch.p) VX_HIP(hipMalloc(&scratch.p, 4 * words));
    inverse_passes(static_cast<uint32_t*>(scratch.p), tw, s, bt, p);
    return true;
  });
  inverse_passes(data, tw, s, bt, plan);
}

void ntt_inverse_cols(uint32_t* const* d_cols, uint32_t n_cols, const TwiddleRef& tw, hipStream_t s) {
  require(d_cols != nullptr && tw.log_n >= 1 && tw.log_n <= 31, "ntt_inverse_cols: bad arguments");
  inverse_passes(nullptr, tw, s, Batch{n_cols, 0, 0, d_cols, nullptr},
                 known_ntt_plan(NttKey{false, tw.log_n, tw.log_n, n_cols, tw.stwo_circle != 0}));
}

// ---------------------------------------------------------------------------
// Uneven batched LDE: columns of log sizes 14..24 (interpolate on their trace
// domain, evaluate on their 2^blowup-times larger eval domain) in four
// work-list launches, whatever the number and sizes of the columns:
//   0 inverse contiguous 14-bit pass (in place on the coefficients)
//   1 inverse strided pass over bits [14, L) (columns with L > 14), 1/n
//   2 forward strided pass over bits [14, L) of the evaluation, reading
//     coefficient i & (n - 1) (the replication layers are never executed)
//   3 forward contiguous 14-bit pass
// The per-column passes are exactly those of ntt_inverse / ntt_forward, so
// every column's result equals the single-column transform.
size_t lde_batch_job_bytes() { return sizeof(RadixJob); }

void lde_batch_build(const LdeBatchCol* cols, uint32_t n, uint32_t blowup_bits, std::vector<uint8_t>& out,
                     LdeBatchStages& st) {
  std::vector<RadixJob> stage[4];
  for (uint32_t c = 0; c < n; c++) {
    const LdeBatchCol& C = cols[c];
    const uint32_t L = C.log_n, E = L + blowup_bits;
    require(L >= RX_E_BITS && L <= 24 && E <= 31 && C.itw.log_n == L && C.etw.log_n == E && C.coef && C.eval &&
                !C.itw.stwo_circle && !C.etw.stwo_circle,
            "lde_batch: columns need 14 <= log_n <= 24 and matching twiddles");
    require(aligned16(C.coef) && aligned16(C.eval) && (!C.in || aligned16(C.in)),
            "lde_batch: columns must be 16-byte aligned");
    const uint32_t e = (30u * L) % 31u;
    const uint32_t inv_n = e == 0 ? 1u : (1u << e);
    const uint64_t nmask = (1ull << L) - 1, Emask = (1ull << E) - 1;
    const uint32_t sb = L - RX_E_BITS;  // strided bits
    stage[0].push_back({C.coef, C.in ? C.in : C.coef, 0,
                        make_radix_args(false, C.itw, 0, RX_E_BITS, 0, nmask, sb == 0 ? inv_n : 1u)});
    if (sb) {
      require(rx_strided_ok(RX_E_BITS, sb, RX_E_BITS - sb, nmask), "lde_batch: strided pass shape");
      stage[1].push_back({C.coef, C.coef, 0, make_radix_args(false, C.itw, RX_E_BITS, sb, RX_E_BITS - sb, nmask, inv_n)});
      stage[2].push_back({C.eval, C.coef, 0, make_radix_args(true, C.etw, RX_E_BITS, sb, RX_E_BITS - sb, nmask, 1u)});
      stage[3].push_back({C.eval, C.eval, 0, make_radix_args(true, C.etw, 0, RX_E_BITS, 0, Emask, 1u)});
    } else {
      stage[3].push_back({C.eval, C.coef, 0, make_radix_args(true, C.etw, 0, RX_E_BITS, 0, nmask, 1u)});
    }
  }
  // tile counts: stages 0, 1 run on the 2^L-element transform, 2, 3 on 2^E
  out.clear();
  for (int k = 0; k < 4; k++) {
    uint64_t tiles = 0;
    uint32_t idx = 0;
    for (uint32_t c = 0; c < n; c++) {
      const uint32_t L = cols[c].log_n, E = L + blowup_bits;
      if (!(k == 0 || k == 3 || L > RX_E_BITS)) continue;
      RadixJob& j = stage[k][idx++];
      j.first_tile = tiles;
      tiles += 1ull << ((k < 2 ? L : E) - RX_E_BITS);
    }
    st.count[k] = (uint32_t)stage[k].size();
    st.tiles[k] = tiles;
    st.offset[k] = out.size();
    const uint8_t* p = reinterpret_cast<const uint8_t*>(stage[k].data());
    out.insert(out.end(), p, p + stage[k].size() * sizeof(RadixJob));
  }
}

void lde_batch_launch(const uint8_t* d_jobs, const LdeBatchStages& st, hipStream_t s) {
  for (int k = 0; k < 4; k++) {
    if (!st.count[k]) continue;
    require(st.tiles[k] < (1ull << 31), "lde_batch: too many tiles");
    const RadixJob* jobs = reinterpret_cast<const RadixJob*>(d_jobs + st.offset[k]);
    // stages 0 and 3 are the contiguous 14-bit passes of every column,
    // stages 1 and 2 strided passes of 1..10 row bits (checked when built)
    if (k == 3)
      hipLaunchKernelGGL((k_ntt_radix_list<true, RX_CONTIG>), dim3((uint32_t)st.tiles[k]), dim3(RX_THREADS), 0, s,
                         jobs, st.count[k]);
    else if (k == 2)
      hipLaunchKernelGGL((k_ntt_radix_list<true, RX_STRIDED>), dim3((uint32_t)st.tiles[k]), dim3(RX_THREADS), 0, s,
                         jobs, st.count[k]);
    else if (k == 0)
      hipLaunchKernelGGL((k_ntt_radix_list<false, RX_CONTIG>), dim3((uint32_t)st.tiles[k]), dim3(RX_THREADS), 0, s,
                         jobs, st.count[k]);
    else
      hipLaunchKernelGGL((k_ntt_radix_list<false, RX_STRIDED>), dim3((uint32_t)st.tiles[k]), dim3(RX_THREADS), 0, s,
                         jobs, st.count[k]);
    VX_HIP(hipGetLastError());
  }
}

// Out-of-place batched bit reverse, dst[rev(i)] = src[i], for columns of
// log size >= 10: a workgroup moves a 32 x 32 tile (the 1024 indices with
// one middle-bit value m) through LDS, reading 32 runs of 32 consecutive
// words and writing 32 runs of 32 consecutive words (rev(a, m, b) =
// (rev5 b, rev m, rev5 a)).
struct BitrevJob {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t log_n;
  uint32_t vec;  // both columns 16-byte aligned: 16 bytes per lane
  uint64_t first_tile;
};
// Columns of log size >= 12 move in 64 x 64 tiles (256-byte runs on both
// sides), smaller ones (10, 11) in 32 x 32 tiles. (128 x 128 tiles, 512-byte
// runs at two workgroups per CU, took the 96-column bit reverse 0.59 ->
// 0.61 ms.)
static constexpr uint32_t BR_BIG = 6, BR_SMALL = 5, BR_MIN_LOG = 2 * BR_SMALL;
static constexpr uint32_t BR_WGS = 2048;  // persistent workgroups (8 per CU)

__host__ __device__ __forceinline__ uint32_t br_tbits(uint32_t log_n) {
  return log_n >= 2 * BR_BIG ? BR_BIG : BR_SMALL;
}

// One tile m of column J: index i = (a << (E - T)) | (m << T) | b maps to
// rev(i) = (rev_T(b) << (E - T)) | (rev(m) << T) | rev_T(a).
template <uint32_t T>
__device__ __forceinline__ void br_tile(const uint32_t* src, uint32_t* dst, uint32_t E, uint64_t m,
                                        uint32_t (*tile)[(1u << BR_BIG) + 1]) {
  constexpr uint32_t W = 1u << T, ROWS = 256 / W;
  const uint32_t mb = E - 2 * T;
  const uint64_t rm = mb ? (__brevll(m) >> (64 - mb)) : 0;
  const uint32_t t = threadIdx.x;
  // row r of a lane is `ROWS r` rows below its first: a uniform step
  const uint64_t step = (uint64_t)ROWS << (E - T);
  const uint32_t* sp = src + (((uint64_t)(t / W) << (E - T)) | (m << T) | (t % W));
#pragma unroll
  for (uint32_t r = 0; r < W / ROWS; r++) tile[(t / W) + ROWS * r][t % W] = sp[r * step];
  __syncthreads();
  uint32_t* dp = dst + (((uint64_t)(t / W) << (E - T)) | (rm << T) | (t % W));
  const uint32_t a = __brev(t % W) >> (32 - T);
#pragma unroll
  for (uint32_t r = 0; r < W / ROWS; r++) {
    const uint32_t b = __brev((t / W) + ROWS * r) >> (32 - T);
    dp[r * step] = tile[a][b];
  }
  __syncthreads();  // the next tile's writes must not overtake these reads
}

// The same tile with 16 bytes per lane on both sides: W / 4 lanes per row,
// and a lane writes destination words rev(a) = 4u .. 4u + 3 of its row,
// taken from four tile rows a. Split into the loads of a tile (into
// registers) and the rest, so that the loop below can load the next tile of
// a column while this one goes through LDS and out.
constexpr uint32_t BR_VEC_LOADS = 4;  // 16-byte loads per lane of a 64 x 64 tile
template <uint32_t T>
__device__ __forceinline__ void br_load_vec(const uint32_t* src, uint32_t E, uint64_t m, uint4 (&v)[BR_VEC_LOADS]) {
  constexpr uint32_t W = 1u << T, LPR = W / 4, ROWS = 256 / LPR;
  const uint32_t t = threadIdx.x, c4 = 4 * (t % LPR);
  const uint64_t step = (uint64_t)ROWS << (E - T);
  const uint32_t* sp = src + (((uint64_t)(t / LPR) << (E - T)) | (m << T) | c4);
#pragma unroll
  for (uint32_t r = 0; r < W / ROWS; r++) v[r] = *reinterpret_cast<const uint4*>(sp + r * step);
}
template <uint32_t T>
__device__ __forceinline__ void br_store_vec(uint32_t* dst, uint32_t E, uint64_t m, const uint4 (&v)[BR_VEC_LOADS],
                                             uint32_t (*tile)[(1u << BR_BIG) + 1]) {
  constexpr uint32_t W = 1u << T, LPR = W / 4, ROWS = 256 / LPR;
  const uint32_t mb = E - 2 * T;
  const uint64_t rm = mb ? (__brevll(m) >> (64 - mb)) : 0;
  const uint32_t t = threadIdx.x, c4 = 4 * (t % LPR);
  const uint64_t step = (uint64_t)ROWS << (E - T);
#pragma unroll
  for (uint32_t r = 0; r < W / ROWS; r++) {
    uint32_t* row = tile[(t / LPR) + ROWS * r] + c4;
    row[0] = v[r].x;
    row[1] = v[r].y;
    row[2] = v[r].z;
    row[3] = v[r].w;
  }
  __syncthreads();
  uint32_t* dp = dst + (((uint64_t)(t / LPR) << (E - T)) | (rm << T) | c4);
  uint32_t a[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) a[k] = __brev(c4 + k) >> (32 - T);
#pragma unroll
  for (uint32_t r = 0; r < W / ROWS; r++) {
    const uint32_t b = __brev((t / LPR) + ROWS * r) >> (32 - T);
    *reinterpret_cast<uint4*>(dp + r * step) = make_uint4(tile[a[0]][b], tile[a[1]][b], tile[a[2]][b], tile[a[3]][b]);
  }
  __syncthreads();  // the next tile's writes must not overtake these reads
}

// Workgroup w runs the contiguous tiles [w chunk, (w + 1) chunk) of the
// whole list, so its job changes rarely and is found by one binary search
// and then a forward scan (scalar loads). Within a run of 64 x 64 tiles of a
// 16-byte-aligned column the next tile's loads are issued before this tile
// is written out.
__global__ __launch_bounds__(256) void k_bitrev_list(const BitrevJob* jobs, uint32_t n_jobs, uint64_t total,
                                                     uint64_t chunk) {
  __shared__ uint32_t tile[1u << BR_BIG][(1u << BR_BIG) + 1];
  const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
  if (b0 >= total) return;
  const uint64_t b1 = b0 + chunk < total ? b0 + chunk : total;
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ld_uniform(&jobs[mid].first_tile) <= b0) lo = mid;
    else hi = mid;
  }
  uint64_t b = b0;
  while (b < b1) {
    while (lo + 1 < n_jobs && ld_uniform(&jobs[lo + 1].first_tile) <= b) lo++;
    uniform_t<BitrevJob>& J = as_uniform(jobs + lo);
    const uint32_t E = J.log_n;
    // this workgroup's run of tiles in column J
    const uint64_t end = lo + 1 < n_jobs ? ld_uniform(&jobs[lo + 1].first_tile) : total;
    const uint64_t run_end = end < b1 ? end : b1;
    if (J.vec && E >= 2 * BR_BIG) {
      // software pipeline over the run: the next tile's loads are issued
      // (unconditionally: the last tile reloads itself) before this tile
      // goes through LDS and out
      uint4 cur[BR_VEC_LOADS], nxt[BR_VEC_LOADS];
      br_load_vec<BR_BIG>(J.src, E, b - J.first_tile, cur);
      for (; b < run_end; b++) {
        const uint64_t m = b - J.first_tile;
        br_load_vec<BR_BIG>(J.src, E, b + 1 < run_end ? m + 1 : m, nxt);
        br_store_vec<BR_BIG>(J.dst, E, m, cur, tile);
#pragma unroll
        for (uint32_t r = 0; r < BR_VEC_LOADS; r++) cur[r] = nxt[r];
      }
    } else {
      for (; b < run_end; b++) {
        const uint64_t m = b - J.first_tile;
        if (J.vec) {
          uint4 v[BR_VEC_LOADS];
          br_load_vec<BR_SMALL>(J.src, E, m, v);
          br_store_vec<BR_SMALL>(J.dst, E, m, v, tile);
        } else if (E >= 2 * BR_BIG) {
          br_tile<BR_BIG>(J.src, J.dst, E, m, tile);
        } else {
          br_tile<BR_SMALL>(J.src, J.dst, E, m, tile);
        }
      }
    }
  }
}

void bitrev_batch_build(const uint32_t* const* src, uint32_t* const* dst, const uint32_t* log_n, uint32_t n,
                        std::vector<uint8_t>& out, uint32_t& count, uint64_t& tiles) {
  std::vector<BitrevJob> jobs;
  tiles = 0;
  for (uint32_t c = 0; c < n; c++) {
    require(log_n[c] >= BR_MIN_LOG && log_n[c] <= 31 && src[c] && dst[c] && src[c] != dst[c],
            "bitrev_batch: out-of-place columns of log size >= 10");
    const bool vec = ((reinterpret_cast<uintptr_t>(src[c]) | reinterpret_cast<uintptr_t>(dst[c])) & 15u) == 0;
    jobs.push_back({src[c], dst[c], log_n[c], vec ? 1u : 0u, tiles});
    tiles += 1ull << (log_n[c] - 2 * br_tbits(log_n[c]));
  }
  count = n;
  out.assign(reinterpret_cast<const uint8_t*>(jobs.data()),
             reinterpret_cast<const uint8_t*>(jobs.data()) + jobs.size() * sizeof(BitrevJob));
}

void bitrev_batch_launch(const uint8_t* d_jobs, uint32_t count, uint64_t tiles, hipStream_t s) {
  if (!count || !tiles) return;
  const uint64_t chunk = (tiles + BR_WGS - 1) / BR_WGS;
  const uint32_t grid = (uint32_t)((tiles + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_bitrev_list, dim3(grid), dim3(256), 0, s, reinterpret_cast<const BitrevJob*>(d_jobs), count,
                     tiles, chunk);
  VX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Single reference layer (cuda/circle_ntt.cu:89-117), for the layer-wise API.
__global__ void k_ntt_layer(uint32_t* data, const uint32_t* tw, uint32_t li, uint64_t half_n, int fwd) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half_n) return;
  uint64_t stride = 1ull << li;
  uint64_t h = t >> li, l = t & (stride - 1);
  uint64_t i0 = (h << (li + 1)) + l, i1 = i0 + stride;
  uint32_t v0 = data[i0], v1 = data[i1], w = tw[h];
  if (fwd) {
    uint32_t m = m31_mul(v1, w);
    data[i0] = m31_add(v0, m);
    data[i1] = m31_sub(v0, m);
  } else {
    data[i0] = m31_add(v0, v1);
    data[i1] = m31_mul(m31_sub(v0, v1), w);
  }
}

void ntt_layer(uint32_t* data, const uint32_t* tw, uint32_t li, uint64_t n, bool forward, hipStream_t s) {
  require(n >= 2 && (n & (n - 1)) == 0 && (1ull << (li + 1)) <= n, "ntt_layer: bad shape");
  uint64_t half = n / 2;
  hipLaunchKernelGGL(k_ntt_layer, dim3((uint32_t)((half + 255) / 256)), dim3(256), 0, s, data, tw, li, half,
                     forward ? 1 : 0);
  VX_HIP(hipGetLastError());
}

__global__ void k_scale(uint32_t* d, uint32_t sc, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = m31_mul(d[i], sc);
}
void m31_scale(uint32_t* d, uint32_t sc, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scale, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, d, sc, n);
  VX_HIP(hipGetLastError());
}

// In-place i <-> bitrev(i) swap (cuda/circle_ntt.cu:234-244).
__global__ void k_bitrev_m31(uint32_t* d, uint32_t log_n, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t j = log_n ? (__brevll(i) >> (64 - log_n)) : 0;
  if (i < j) {
    uint32_t t = d[i];
    d[i] = d[j];
    d[j] = t;
  }
}
void bit_reverse_m31(uint32_t* d, uint32_t log_n, hipStream_t s) {
  require(log_n <= 31, "bit_reverse: log_n too large");
  uint64_t n = 1ull << log_n;
  hipLaunchKernelGGL(k_bitrev_m31, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, d, log_n, n);
  VX_HIP(hipGetLastError());
}

// Out-of-place QM31 (4 words per element, AoS) bit reverse
// (cuda/bit_reverse_wide.cu:21-52).
__global__ void k_bitrev_qm31(const uint4* in, uint4* out, uint32_t log_n, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t j = log_n ? (__brevll(i) >> (64 - log_n)) : 0;
  out[j] = in[i];
}
void bit_reverse_qm31(const uint32_t* in, uint32_t* out, uint32_t log_n, hipStream_t s) {
  require(log_n <= 31 && in != out, "bit_reverse_qm31: bad arguments");
  uint64_t n = 1ull << log_n;
  hipLaunchKernelGGL(k_bitrev_qm31, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, (const uint4*)in,
                     (uint4*)out, log_n, n);
  VX_HIP(hipGetLastError());
}

// dst[i] = i < src_n ? src[i] : 0 (cuda/constraint.cu:81-90).
__global__ void k_zero_pad(const uint32_t* src, uint32_t* dst, uint64_t src_n, uint64_t dst_n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dst_n) dst[i] = i < src_n ? src[i] : 0u;
}
void zero_pad(const uint32_t* src, uint32_t* dst, uint64_t src_n, uint64_t dst_n, hipStream_t s) {
  if (dst_n == 0) return;
  hipLaunchKernelGGL(k_zero_pad, dim3((uint32_t)((dst_n + 255) / 256)), dim3(256), 0, s, src, dst, src_n,
                     dst_n);
  VX_HIP(hipGetLastError());
}

}  // namespace vx
