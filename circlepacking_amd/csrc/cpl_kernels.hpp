// Worklist layout and device kernels of the plain FP64 prover.
#pragma once

#include "cpl_lane.hpp"
#include "cpl_mlp.hpp"
#include "cpl.h"

namespace cpl {

// ---- worklist (SoA, one region per family) ----------------------------------
// e[k][i]: the 8 interval endpoints of item i (la.lb, la.ub, r1.lb, r1.ub,
// r2.lb, r2.ub, p3.lb, p3.ub), each a coalesced array of doubles;
// root[i]: root cell index; la_idx[i]: index of the item's la cell at its own
// la depth (key of the lower_bound_radius_single_placement table);
// meta[i]: depth << 16 | splits (4 bits per axis: bisections so far).
struct Worklist {
  double* e[8];
  uint32_t* root;
  uint32_t* la_idx;
  uint32_t* meta;
};

// Per (round, family) statistics are kept in STAT_SLOTS copies (one 128-byte
// line each) and each wave adds into the copy of its index: tens of thousands
// of waves adding to one line serialise in L2 and cost as much as the
// evaluation itself.  The copies are summed on the device after the run.
constexpr int STAT_SLOTS = 64;

struct RoundStats {
  unsigned long long evaluations;
  unsigned long long verdicts[8];
  unsigned long long digest;
  unsigned long long vdigest;
  unsigned long long certified;
  unsigned long long unresolved;
  unsigned long long fp32_settled;  // items the FP32 prefilter proved
  unsigned long long fp32_tried;    // items the FP32 prefilter ran on
  unsigned long long pad;
};
static_assert(sizeof(RoundStats) == 16 * sizeof(unsigned long long), "16 words, RoundStats word order used by kernels");

struct RefineParams {
  int32_t grid[4];
  double box_lb[4], box_ub[4];
  int32_t clip;
  int32_t max_depth;
  int32_t split_mode;
  uint32_t split_axes;
  int32_t n_children;     // children per split (2, or 2^popcount(split_axes) in mode 1)
  int32_t nfam;           // families in the run
  int32_t fam_slot;       // slot of this launch's family among them
  int32_t depth;          // round index
  uint64_t n_roots;
  const double* la_table; // per la cell: lb, ub, lbrsp; level d starts at grid[0]*(2^d - 1)
  uint32_t* root_flags;   // bytes, (root * nfam + slot), accessed as words
  RoundStats* stats;      // [CPL_MAX_ROUNDS][CPL_NUM_FAMILIES]
  Worklist out;           // next-round worklist region of this family
  uint32_t* out_count;
  uint64_t out_cap;
  uint32_t* overflow;
  const uint32_t* in_count;  // device-side item count of the input (graph mode), or null
  int32_t digest;            // 1: accumulate the per-item digests
  uint32_t prefilter_depths; // extended kernels: bit d enables the FP32 prefilter at depth d
  const MlpPacked* mlp;      // extended kernels: learned proposal of this family, or null
  float mlp_threshold;       // gate: the prefilter runs when output 0 exceeds this logit
  int32_t learned_split;     // 1: mode-0 splits take the proposal's best-scored eligible axis
  const uint8_t* prop;       // extended round kernels: proposal byte per input item (k_propose), or null
  uint32_t* zero_count;      // cleared by the round kernel (the next round's output counter), or null
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Same definition as the oracle's worklist digest.
__device__ __forceinline__ uint64_t item_hash(const double e[8], uint32_t root, int family, uint32_t meta) {
  uint64_t h = 0x243F6A8885A308D3ull;
#pragma unroll
  for (int k = 0; k < 8; ++k) h = mix64(h ^ dbits(e[k]));
  uint64_t m = ((uint64_t)root << 32) | ((uint64_t)family << 24) | (uint64_t)meta;
  return mix64(h ^ m);
}

__device__ __forceinline__ int axis_splits(uint32_t meta, int a) { return (meta >> (4 * a)) & 15; }

}  // namespace cpl
