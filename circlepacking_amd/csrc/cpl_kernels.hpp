// Worklist layout and device kernels of the plain FP64 prover.
#pragma once

#include "cpl_lane.hpp"
#include "cpl_mlp.hpp"
#include "cpl.h"

namespace cpl {

// ---- worklist (SoA, one region per family) ----------------------------------
// e(k, i): the 8 interval endpoints of item i (la.lb, la.ub, r1.lb, r1.ub,
// r2.lb, r2.ub, p3.lb, p3.ub), each a coalesced array of doubles;
// root(i): root cell index; la_idx(i): index of the item's la cell at its own
// la depth (key of the lower_bound_radius_single_placement table);
// meta(i): depth << 16 | splits (4 bits per axis: bisections so far).
// A region is addressed by one base pointer per kind and its capacity rather
// than eleven array pointers: the round kernels address an input and an output
// region in their loop, and every pointer the compiler keeps live across the
// evaluation takes scalar registers from it.
struct Worklist {
  double* ev;    // e(k, i) = ev[k * cap + i]
  uint32_t* uv;  // root, la_idx, meta: uv[i], uv[cap + i], uv[2 * cap + i]
  uint64_t cap;
  __device__ __forceinline__ double& e(int k, uint64_t i) const { return ev[(uint64_t)k * cap + i]; }
  __device__ __forceinline__ uint32_t& root(uint64_t i) const { return uv[i]; }
  __device__ __forceinline__ uint32_t& la_idx(uint64_t i) const { return uv[cap + i]; }
  __device__ __forceinline__ uint32_t& meta(uint64_t i) const { return uv[2 * cap + i]; }
};

// ---- segmented regions -------------------------------------------------------
// A family's region of kSegs * seg_cap items is kSegs segments; segment s holds
// the items [s * seg_cap, s * seg_cap + min(count[s], seg_cap)).  Each segment
// count is a 64-bit word in its own 128-byte line.  A wave appends to one
// segment with one atomic and moves on to the next segment that is not full
// when its segment fills; waves start on different segments.  (Returning
// device-scope atomics on a single address serialise: 10,800 waves appending
// to one counter took 127 us on an MI355X, to 64 counters 7.7 us.)  Readers
// address item i of the region (0 <= i < sum of the clamped counts) through the
// prefix sums of the counts.  Every position below min(count, seg_cap) is
// written: an append that runs past seg_cap writes the part that fits and
// places the rest in the next segment.
constexpr int kSegs = 64;
constexpr int kSegStride = 16;  // counter words (u64) between segment counts

struct SegOut {
  unsigned long long* cnt;  // kSegs counts, kSegStride words apart
  uint32_t seg_cap;
  uint32_t* overflow;       // set when every segment is full
};

// Wave-collective (all 64 lanes active): inclusive prefix over the segments of
// the clamped counts, segment s in lane s; returns the item total.
__device__ __forceinline__ uint64_t seg_scan(const unsigned long long* cnt, uint32_t seg_cap, uint64_t& incl) {
  const int lane = threadIdx.x & 63;
  uint64_t v = cnt[lane * kSegStride];
  v = v < seg_cap ? v : seg_cap;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  incl = v;
  return __shfl(v, 63, 64);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Wave-collective: storage index of item i (i < total; any value on lanes
// whose result is unused) from the prefix `incl` of seg_scan.
__device__ __forceinline__ uint64_t seg_locate(uint64_t i, uint64_t incl, uint32_t seg_cap) {
  int s = 0;  // number of segments whose inclusive prefix is <= i (at most 63 for i < total)
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    const uint64_t v = __shfl(incl, s + step - 1, 64);
    if (v <= i) s += step;
  }
  const uint64_t before = __shfl(incl, (s + 63) & 63, 64);
  return (uint64_t)s * seg_cap + (i - (s ? before : 0ull));
}

// Same, for a wave whose lanes hold the consecutive items i = i0 + lane: the
// segment of i0 comes from one ballot, and when all 64 items lie in it (the
// usual case) no per-lane search runs.
__device__ __forceinline__ uint64_t seg_locate_run(uint64_t i0, uint64_t i, uint64_t incl, uint32_t seg_cap) {
  const int s0 = __popcll(__ballot(incl <= i0));  // 64 only when i0 >= total (no valid lane)
  const uint64_t end0 = readlane64(incl, s0 & 63);
  const uint64_t before0 = s0 ? readlane64(incl, (s0 - 1) & 63) : 0ull;
  if (s0 < 64 && i0 + 63 < end0) return (uint64_t)s0 * seg_cap + (i - before0);
  return seg_locate(i, incl, seg_cap);
}

// Per-wave append state (wave-uniform).
struct SegCursor {
  uint64_t full;  // segments this wave found full
  int next;       // segment to try first
};

__device__ __forceinline__ SegCursor seg_cursor() {
  const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  return SegCursor{0ull, wave & (kSegs - 1)};
}

// One reservation attempt: `need` slots for this lane.  Lane 0 issues the
// atomic; its result is first read by seg_commit, so work placed between the
// two overlaps the atomic's round trip.
struct SegTicket {
  int s;            // segment, or -1 (nothing to place, or every segment full)
  uint64_t b;       // lane 0: segment count before the append
  uint64_t excl;    // slots of the lower lanes
  uint64_t total;   // slots of the wave
  uint32_t need;
};

__device__ __forceinline__ int seg_pick(const SegCursor& c) {
  const uint64_t avail = ~c.full;
  if (!avail) return -1;
  const int st = c.next;
  const uint64_t rot = st ? ((avail >> st) | (avail << (64 - st))) : avail;
  return (st + __ffsll((unsigned long long)rot) - 1) & (kSegs - 1);
}

// First attempt, every lane with `want` needing the same `nc` slots (ranks
// from a ballot).
__device__ __forceinline__ SegTicket seg_reserve(const SegOut& o, const SegCursor& c, bool want, uint32_t nc) {
  const uint64_t mask = __ballot(want);
  SegTicket t;
  t.need = want ? nc : 0u;
  t.total = (uint64_t)__popcll(mask) * nc;
  t.excl = (uint64_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u)) * nc;
  t.s = t.total ? seg_pick(c) : -1;
  t.b = 0;
  if (t.s >= 0 && (threadIdx.x & 63) == 0) t.b = atomicAdd(o.cnt + t.s * kSegStride, (unsigned long long)t.total);
  return t;
}

// Later attempts, per-lane needs (prefix scan over the wave).
__device__ __forceinline__ SegTicket seg_reserve_any(const SegOut& o, const SegCursor& c, uint32_t need) {
  const int lane = threadIdx.x & 63;
  SegTicket t;
  t.need = need;
  uint64_t incl = need;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  t.total = readlane64(incl, 63);
  t.excl = incl - need;
  t.s = t.total ? seg_pick(c) : -1;
  t.b = 0;
  if (t.s >= 0 && lane == 0) t.b = atomicAdd(o.cnt + t.s * kSegStride, (unsigned long long)t.total);
  return t;
}

// Completes a reservation: emit(k, index) writes this lane's k-th slot at
// storage index `index`; slots that do not fit go to the next segments.  Every
// attempt that leaves a slot unplaced marks its segment full, so the loop runs
// at most kSegs + 1 times; when every segment is full the overflow word is set.
template <class Emit>
__device__ __forceinline__ void seg_commit(const SegOut& o, SegCursor& c, SegTicket t, Emit&& emit) {
  uint32_t done = 0;
  while (t.total) {
    if (t.s < 0) {
      if (t.need) atomicOr(o.overflow, 1u);
      return;
    }
    const uint64_t b = readlane64(t.b, 0);
    const uint64_t p = b + t.excl;
    const uint32_t fit = p >= o.seg_cap ? 0u : (uint32_t)min((uint64_t)t.need, (uint64_t)o.seg_cap - p);
    const uint64_t at = (uint64_t)t.s * o.seg_cap + p;
    for (uint32_t k = 0; k < fit; ++k) emit(done + k, at + k);
    done += fit;
    c.next = t.s;
    if (b + t.total >= o.seg_cap) c.full |= 1ull << t.s;
    const uint32_t rest = t.need - fit;
    if (!__ballot(rest != 0u)) return;
    t = seg_reserve_any(o, c, rest);
  }
}

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
  unsigned long long live_items;    // round 0 of the graph form: root items the triple screening passed on
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
  unsigned long long* out_seg;  // its kSegs segment counts
  uint32_t seg_cap;             // items per segment (input and output regions)
  uint32_t* overflow;
  const unsigned long long* in_seg;  // segment counts of the input region, or null (items 0..n-1 of `in`)
  int32_t digest;            // 1: accumulate the per-item digests
  uint32_t prefilter_depths; // extended kernels: bit d enables the FP32 prefilter at depth d
  const MlpPacked* mlp;      // extended kernels: learned proposal of this family, or null
  float mlp_threshold;       // gate: the prefilter runs when output 0 exceeds this logit
  int32_t learned_split;     // 1: mode-0 splits take the proposal's best-scored eligible axis
  int32_t learned_order;     // 1: a wave runs first the check its proposals predict to fail first
  const uint8_t* prop;       // extended round kernels: proposal byte per input item (k_propose), or null
  unsigned long long* zero_seg;  // segment counts cleared by the round kernel (the next round's output), or null
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
