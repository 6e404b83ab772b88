// libcpl_hip.so: gfx950 kernels of the lane-packing prover and the C ABI of
// include/cpl.h.
//
// Kernels:
//   k_la_table        la cells at every depth + lower_bound_radius_single_placement
//   k_round<F>        one refinement round of family F over a worklist (or the
//                     implicit root grid): evaluate, count, compact critical
//                     cells' children into the next worklist by wave ballot +
//                     one atomic per wave
//   k_eval_cells<F>   per-cell verdicts of caller-given cells
//   k_body<F>         reference kernel body on explicit (la, r1_, r2_, r3)
//   k_mode_r<F>       one reference-grid launch (the reference's kernel contract)
//   k_persist<F>      one persistent launch per family draining a mixed-depth
//                     worklist from wave-private LDS stacks
//   k_debug_prim, k_debug_geom, k_debug_prefilter
//                     primitives, geometry and FP32-prefilter results for the
//                     parity tests
// EXT template variants run the FP32 outward-rounded prefilter before the FP64
// criteria; k_propose writes the learned proposal (BF16 MFMA) of each round's
// cells ahead of the round.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cpl_host.hpp"
#include "cpl_kernels.hpp"
#include "cpl_trig.hpp"

using namespace cpl;

// ============================================================================
// Device kernels
// ============================================================================

__global__ void k_debug_prim(int op, const double* a, const double* b, double* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = a[i], y = b ? b[i] : 0.0, r = 0.0;
  switch (op) {
    case 0: r = add_rd(x, y); break;
    case 1: r = add_ru(x, y); break;
    case 2: r = sub_rd(x, y); break;
    case 3: r = sub_ru(x, y); break;
    case 4: r = mul_rd(x, y); break;
    case 5: r = mul_ru(x, y); break;
    case 6: r = div_rd(x, y); break;
    case 7: r = div_ru(x, y); break;
    case 8: r = sqrt_rd(x); break;
    case 9: r = sqrt_ru(x); break;
    case 10: r = asin_rn(x); break;
    case 11: r = dasin_rd(x); break;
    case 12: r = dasin_ru(x); break;
    case 13: r = lower_bound_radius_single_placement(x); break;
    case 14: r = x / y; break;
    case 15: r = __builtin_sqrt(x); break;
    case 16: r = add_rd((float)x, (float)y); break;
    case 17: r = add_ru((float)x, (float)y); break;
    case 18: r = add_rd((float)x, -(float)y); break;
    case 19: r = add_ru((float)x, -(float)y); break;
    case 20: r = mul_rd((float)x, (float)y); break;
    case 21: r = mul_ru((float)x, (float)y); break;
    case 22: r = div_rd((float)x, (float)y); break;
    case 23: r = div_ru((float)x, (float)y); break;
    case 24: r = asin32_rn((float)x); break;
    case 25: r = dasin_rd((float)x); break;
    case 26: r = dasin_ru((float)x); break;
    case 27: r = f32_down(x); break;
    case 28: r = f32_up(x); break;
    case 29: r = sin_rn(x); break;
    case 30: r = cos_rn(x); break;
    case 31: r = acos_rn(x); break;
    case 32: r = dsin_rd(x); break;
    case 33: r = dsin_ru(x); break;
    case 34: r = dcos_rd(x); break;
    case 35: r = dcos_ru(x); break;
    case 36: r = dacos_rd(x); break;
    case 37: r = dacos_ru(x); break;
    case 38: r = dfrom(ordered_of(x)); break;         // the u64 image carried in the bits
    case 39: r = double_of_ordered(dbits(x)); break;  // inverse of the image in x's bits
    case 40: r = __builtin_amdgcn_rcp(x); break;      // the hardware reciprocal estimate
    case 41: r = __builtin_amdgcn_rsq(x); break;      // the hardware reciprocal square root estimate
    default: r = __builtin_nan("");
  }
  out[i] = r;
}

// Elementwise interval operations (the algcuda::Interval<double> surface).
__global__ void k_interval_op(int op, const IV* a, const IV* b, IV* out, int* trap, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const IV x = a[i];
  const IV y = b ? b[i] : IV(0.0, 0.0);
  IV r(0.0, 0.0);
  int t = 0;
  switch (op) {
    case 0: r = x + y; break;
    case 1: r = x - y; break;
    case 2: r = x * y; break;
    case 3: r = x / y; break;
    case 4: r = sqrt_iv(x, t); break;
    case 5: r = x.square(); break;
    case 6: r = x.abs(); break;
    case 7: r = reciprocal_iv(x); break;
    case 8: r = div_scalar(x, y.lb); break;
    case 9: r = IV(x.lb, x.lb) / y; break;
    case 10: r = sin_iv(x); break;
    case 11: r = cos_iv(x); break;
    case 12: r = asin_iv(x, t); break;
    case 13: r = acos_iv(x, t); break;
    case 14: r = x * y.lb; break;
    default: r = neg(x); break;
  }
  out[i] = r;
  trap[i] = t;
}

// Geometry functions of the criteria on interval records (la, ra, rb), in
// FP64 (T = double) or in the FP32 prefilter precision (T = float, inputs
// converted outward).  fn: 0 inner_inner 1 outer_inner 2 inner_outer
// 3 outer_outer(ra, rb) 4 semicircle_outer(ra) 5 semicircle_inner(la, ra)
// 6 compute_collision_angle(lc = la, sc = ra, ab = rb), 7 / 8 the same with
// only the lower / upper endpoint computed.
template <class T>
__device__ __forceinline__ IVT<T> conv(IV x);
template <>
__device__ __forceinline__ IV conv<double>(IV x) { return x; }
template <>
__device__ __forceinline__ IVf conv<float>(IV x) { return to_f32(x); }

template <class T>
__global__ void k_debug_geom(int fn, const double* in, double* out, int* trap, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = in + 6 * i;
  IVT<T> la = conv<T>(IV(p[0], p[1])), ra = conv<T>(IV(p[2], p[3])), rb = conv<T>(IV(p[4], p[5]));
  IVT<T> r(T(0), T(0));
  int t = 0;
  switch (fn) {
    case 0: r = collision_angle_inner_inner(la, ra, rb, t); break;
    case 1: r = collision_angle_outer_inner(la, ra, rb, t); break;
    case 2: r = collision_angle_inner_outer(la, ra, rb, t); break;
    case 3: r = collision_angle_outer_outer(ra, rb, t); break;
    case 4: r = semicircle_angle_outer(ra); break;
    case 5: r = semicircle_angle_inner(la, ra); break;
    case 6: r = compute_collision_angle<3>(la, ra, rb, t); break;
    case 7: r = compute_collision_angle<1>(la, ra, rb, t); break;
    default: r = compute_collision_angle<2>(la, ra, rb, t); break;
  }
  out[2 * i] = r.lb;
  out[2 * i + 1] = r.ub;
  trap[i] = t;
}

// Per cell: the FP32 criteria result of the tightened cell (0 all hold, k the
// first failing criterion, 16 + trap code; 255 when the FP64 body decides
// before the criteria) and the FP64 verdict byte.
template <int F>
__global__ __launch_bounds__(256) void k_debug_prefilter(const double* cells, size_t n, uint8_t* out32, uint8_t* out64) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* c = cells + 8 * i;
  IV la(c[0], c[1]), r1(c[2], c[3]), r2(c[4], c[5]), r3(c[6], c[7]);
  uint8_t v64 = eval_kernel_body<F>(la, r1, r2, r3);
  if constexpr (is_3d(F)) r2.tighten_lb(r3.lb);
  r1.tighten_lb(r2.lb);
  uint8_t v32 = 255;
  int v = v64 & 7;
  if (v == V_PROVEN || v == V_CRITICAL || v == V_TRAP)
    v32 = (uint8_t)criteria<F, float>(to_f32(la), to_f32(r1), to_f32(r2), to_f32(r3));
  out32[i] = v32;
  out64[i] = v64;
}

// la cells of every depth 0..levels-1: level 0 = get_subinterval(box, j, n0),
// level d+1 = get_subinterval(parent, c, 2).  Entry = (lb, ub, lbrsp(lb)).
__global__ void k_la_table(double box_lb, double box_ub, int n0, int levels, int want_lbrsp, double* table) {
  uint64_t total = (uint64_t)n0 * ((1ull << levels) - 1);
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int d = 0;
  uint64_t off = 0;
  while (i >= off + ((uint64_t)n0 << d)) { off += (uint64_t)n0 << d; ++d; }
  uint64_t j = i - off;
  IV la = get_subinterval(IV(box_lb, box_ub), (int)(j >> d), n0);
  for (int k = d - 1; k >= 0; --k) la = get_subinterval(la, (int)((j >> k) & 1), 2);
  table[3 * i + 0] = la.lb;
  table[3 * i + 1] = la.ub;
  table[3 * i + 2] = want_lbrsp ? lower_bound_radius_single_placement(la.lb) : 0.0;
}

__device__ __forceinline__ uint64_t la_level_offset(int n0, int d) { return (uint64_t)n0 * ((1ull << d) - 1); }

// Statistics copy of (round d, family F) for slot `slot`.
__device__ __forceinline__ RoundStats* stat_slot(const RefineParams& P, int d, int F, uint32_t slot) {
  return P.stats + ((size_t)(d * CPL_NUM_FAMILIES + F) * STAT_SLOTS + (slot & (STAT_SLOTS - 1)));
}
__device__ __forceinline__ uint32_t wave_slot() { return (blockIdx.x * blockDim.x + threadIdx.x) >> 6; }

// Block-wide sum of per-thread counters c[16] (RoundStats word order) added to
// *dst with one atomic per nonzero word: waves reduce by shuffles, wave
// partials meet in LDS.  Called once, at the end of a kernel, by every thread.
__device__ __forceinline__ void block_stats_add(RoundStats* dst, const unsigned long long* c) {
  __shared__ unsigned long long part[16][16];  // [wave][word], blocks of up to 1024 threads
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    unsigned long long v = c[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) part[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    unsigned long long sum = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) sum += part[i][threadIdx.x];
    if (sum) atomicAdd(reinterpret_cast<unsigned long long*>(dst) + threadIdx.x, sum);
  }
}

// Wave totals c[16] (the same value in every lane) added to *dst by lanes
// 0..15, one word each, with no barrier: a wave that finishes early leaves
// at once instead of holding its registers until the block's slowest wave
// reaches a block-wide reduction.
__device__ __forceinline__ void wave_stats_add(RoundStats* dst, const unsigned long long* c) {
  const int lane = threadIdx.x & 63;
  unsigned long long v = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if (lane == k) v = c[k];
  if (lane < 16 && v) atomicAdd(reinterpret_cast<unsigned long long*>(dst) + lane, v);
}

__device__ __forceinline__ void flag_root(uint32_t* flags, uint64_t byte_index, uint32_t bits) {
  atomicOr(flags + (byte_index >> 2), bits << (8 * (byte_index & 3)));
}

// Axis of a mode-0 split: the widest eligible axis in root-extent units
// (smallest grid[a] << splits[a], ties lowest).
__device__ __forceinline__ int widest_axis(const RefineParams& P, uint32_t meta) {
  int best = -1;
  uint64_t bw = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (!(P.split_axes & (1u << a))) continue;
    uint64_t w = (uint64_t)P.grid[a] << axis_splits(meta, a);
    if (best < 0 || w < bw) { best = a; bw = w; }
  }
  return best;
}

// Learned proposal of one wave (all 64 lanes call it) with the network's
// operands already in registers: outputs 0..5 of each lane's cell.  It runs in a
// kernel of its own (k_propose) ahead of each round, so that the evaluation
// kernels keep their register budget.
__device__ __forceinline__ void propose(const MlpBox& B, const MlpRegs& W, bool valid, const double* e, float* y) {
  float f[8];
  if (valid) mlp_features(e, B, f);
  else
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = 0.0f;
  mlp_infer_wave(W, f, y);
}

// Proposal byte of a cell: bit 0 gate (run the FP32 prefilter), bits 1-2 the
// split axis (the eligible axis of highest predicted score y[1 + a]; the
// widest axis wins ties and NaN scores), bit 3 the second check predicted to
// fail first (y[5] > 0), bit 7 set.
__device__ __forceinline__ uint8_t proposal_byte(const RefineParams& P, const float* y, uint32_t meta) {
  int axis = widest_axis(P, meta);
  float ls = y[1 + axis];
#pragma unroll
  for (int a = 0; a < 4; ++a)
    if ((P.split_axes & (1u << a)) && y[1 + a] > ls) { axis = a; ls = y[1 + a]; }
  return (uint8_t)(0x80u | (uint32_t)(axis << 1) | (y[0] > P.mlp_threshold ? 1u : 0u) | (y[5] > 0.0f ? 8u : 0u));
}
// Without a proposal: gate open, widest axis.
constexpr uint8_t kNoProposal = 0x01;

__device__ __forceinline__ int split_axis(const RefineParams& P, uint32_t meta, uint8_t pb) {
  return (P.learned_split && (pb & 0x80)) ? (int)((pb >> 1) & 3) : widest_axis(P, meta);
}

// Proposal bytes of a round's input items (before the round runs on them).
// Each wave loads the network's operands into registers once and then walks
// 64-cell groups with a wave-level grid stride (the operands are 32 KB per
// wave; reloading them per group dominated the kernel's time).
// Two waves per SIMD (at most 256 registers): the MFMA results then go to
// vector registers directly instead of accumulation registers read back one
// by one.
__global__ __launch_bounds__(256, 2) void k_propose(RefineParams P, Worklist in, uint64_t n, uint8_t* out) {
  uint64_t incl = 0;
  if (P.in_seg) n = seg_scan(P.in_seg, P.seg_cap, incl);
  MlpRegs W;
  mlp_load(P.mlp, W);
  const MlpBox B = mlp_box(P.box_lb, P.box_ub);
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  const uint64_t lane = threadIdx.x & 63;
  for (uint64_t grp = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); grp * 64 < n; grp += waves) {
    const uint64_t i = grp * 64 + lane;
    const bool valid = i < n;
    const uint64_t a = P.in_seg ? seg_locate_run(grp * 64, i, incl, P.seg_cap) : i;
    double e[8];
    uint32_t meta = 0;
    if (valid) {
#pragma unroll
      for (int k = 0; k < 8; ++k) e[k] = in.e(k, a);
      meta = in.meta(a);
    }
    float y[6];
    propose(B, W, valid, e, y);
    if (valid) out[i] = proposal_byte(P, y, meta);
  }
}

// One round of family F.  ROOT: items are the implicit root grid (n = n_roots),
// else items come from `in` (n items).
// LB: launch bound (maximum threads per block) the register allocation is made for.
// EXT: extended kernels with the FP32 prefilter (enabled per depth by
// P.prefilter_depths, gated by the proposal byte).  The proposal byte itself
// (P.prop, written by k_propose) is read by every refinement kernel.
// WAVES: the second launch bound, waves per SIMD the registers are budgeted for:
// 3 (<= 168 VGPRs) in every form.  With 4 (<= 128 VGPRs) the families' kernels
// spill 21 to 81 VGPRs per lane and the graph form's proof took 11 % longer;
// with 2 both the graph and the plain form were slower.
template <int F, bool ROOT, int LB = 256, bool EXT = false, int WAVES = 3>
__global__ __launch_bounds__(LB, WAVES) void k_round(RefineParams P, Worklist in, uint64_t n) {
  uint64_t incl = 0;
  if (!ROOT && P.in_seg) n = seg_scan(P.in_seg, P.seg_cap, incl);
  if (P.zero_seg && blockIdx.x == 0 && threadIdx.x < kSegs) P.zero_seg[threadIdx.x * kSegStride] = 0ull;
  const SegOut so{P.out_seg, P.seg_cap, P.overflow};
  SegCursor cur = seg_cursor();
  const int lane = threadIdx.x & 63;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // Wave-uniform counters (kept in scalar registers: counts come from ballots,
  // digests are reduced across the wave each iteration), so no vector
  // registers are held across the evaluation for statistics.  32-bit counts
  // of the six verdicts and of the split items; evaluations, certified and
  // unresolved items follow from them at the end (fewer scalar registers live
  // across the evaluation, which spills scalars into vector lanes otherwise).
  uint32_t c_v[6] = {0, 0, 0, 0, 0, 0}, c_split = 0, c_f32 = 0, c_f32t = 0;
  uint64_t c_dig = 0, c_vdig = 0;

  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    uint64_t i = base + threadIdx.x;
    bool valid = i < n;
    uint64_t a = i;
    if (!ROOT && P.in_seg) a = seg_locate_run(base + (threadIdx.x & ~63u), i, incl, P.seg_cap);
    double e[8];
    uint32_t root = 0, la_idx = 0, meta = 0;
    uint8_t vb = 0;
    bool split = false;
    bool hit = false, tried = false;
    uint64_t h = 0, hv = 0;
    if (valid) {
      if (ROOT) {
        uint64_t r = i;
        int idx[4];
#pragma unroll
        for (int a = 3; a >= 0; --a) { idx[a] = (int)(r % (uint64_t)P.grid[a]); r /= (uint64_t)P.grid[a]; }
        e[0] = P.la_table[3 * idx[0] + 0];
        e[1] = P.la_table[3 * idx[0] + 1];
#pragma unroll
        for (int a = 1; a < 4; ++a) {
          IV s = get_subinterval(IV(P.box_lb[a], P.box_ub[a]), idx[a], P.grid[a]);
          e[2 * a] = s.lb;
          e[2 * a + 1] = s.ub;
        }
        root = (uint32_t)i;
        la_idx = (uint32_t)idx[0];
        meta = 0;
      } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) e[k] = in.e(k, a);
        root = in.root(a);
        la_idx = in.la_idx(a);
        meta = in.meta(a);
      }
    }
    uint8_t pb = kNoProposal;
    if (!ROOT && P.prop && valid) pb = P.prop[i];
    // learned check order: the wave runs the second check first when most of
    // its cells' proposals predict that check to fail first
    bool second_first = false;
    if (!ROOT && P.learned_order) {
      const int voters = __popcll(__ballot(valid && (pb & 0x80)));
      second_first = 2 * __popcll(__ballot(valid && (pb & 0x88) == 0x88)) > voters;
    }
    if (valid) {
      int depth = (int)(meta >> 16);
      double lbrsp = 0.0;
      if (P.clip) lbrsp = P.la_table[3 * (la_level_offset(P.grid[0], axis_splits(meta, 0)) + la_idx) + 2];
      const bool pre = EXT && ((P.prefilter_depths >> depth) & 1u) && (pb & 1);
      vb = eval_synthetic<F, EXT>(IV(e[0], e[1]), IV(e[2], e[3]), IV(e[4], e[5]), IV(e[6], e[7]), P.clip != 0, lbrsp,
                                  pre, &hit, &tried, second_first);
      int v = vb & 7;
      if (P.digest) {
        h = item_hash(e, root, F, meta);
        hv = mix64(h ^ vb);
      }
      uint64_t fbyte = (uint64_t)root * P.nfam + P.fam_slot;
      uint32_t fl = (vb & 8) ? CPL_RF_NOFIT_FLAG : 0;
      if (v == V_CRITICAL && depth < P.max_depth) {
        split = true;
      } else {
        switch (v) {
          case V_PROVEN: fl |= CPL_RF_PROVEN; break;
          case V_EMPTY: fl |= CPL_RF_EMPTY; break;
          case V_OOD: fl |= CPL_RF_OOD; break;
          case V_CRITICAL: fl |= CPL_RF_CRITICAL; break;
          case V_NOFIT: fl |= CPL_RF_NOFIT; break;
          default: fl |= CPL_RF_TRAP; break;
        }
      }
      if (fl && P.root_flags) flag_root(P.root_flags, fbyte, fl);
    }
    // ---- children: one segment append per wave, reserved before the
    // statistics so the atomic's round trip overlaps them ----
    const SegTicket tk = seg_reserve(so, cur, split, (uint32_t)P.n_children);
    {
      const int v = vb & 7;
#pragma unroll
      for (int k = 0; k < 6; ++k) c_v[k] += (uint32_t)__popcll(__ballot(valid && v == k));
      c_split += (uint32_t)__popcll(__ballot(split));
      if (EXT) {
        c_f32 += __popcll(__ballot(hit));
        c_f32t += __popcll(__ballot(tried));
      }
      if (P.digest) {
        uint64_t dg = h, dv = hv;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          dg += __shfl_xor(dg, off, 64);
          dv += __shfl_xor(dv, off, 64);
        }
        c_dig += dg;
        c_vdig += dv;
      }
    }
    if (tk.total) {
      const int depth = (int)(meta >> 16);
      const uint32_t splits = meta & 0xFFFFu;
      if (P.split_mode == 0) {
        const int best = split ? split_axis(P, meta, pb) : 0;
        const uint32_t cmeta = ((uint32_t)(depth + 1) << 16) | (splits + (1u << (4 * best)));
        seg_commit(so, cur, tk, [&](uint32_t c, uint64_t o) {
          const IV sc = get_subinterval(IV(e[2 * best], e[2 * best + 1]), (int)c, 2);
#pragma unroll
          for (int k = 0; k < 8; ++k) P.out.e(k, o) = e[k];
          P.out.e(2 * best, o) = sc.lb;
          P.out.e(2 * best + 1, o) = sc.ub;
          P.out.root(o) = root;
          P.out.la_idx(o) = best == 0 ? la_idx * 2 + c : la_idx;
          P.out.meta(o) = cmeta;
        });
      } else {
        // all eligible axes at once: child c takes bit k of c for the k-th eligible axis
        uint32_t cs = splits;
        for (int ax = 0; ax < 4; ++ax)
          if (P.split_axes & (1u << ax)) cs += (1u << (4 * ax));
        const uint32_t cmeta = ((uint32_t)(depth + 1) << 16) | cs;
        const int nbits = __popc(P.split_axes);
        seg_commit(so, cur, tk, [&](uint32_t c, uint64_t o) {
          double ce[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) ce[k] = e[k];
          uint32_t cla = la_idx;
          int bit = 0;
          for (int ax = 0; ax < 4; ++ax) {
            if (!(P.split_axes & (1u << ax))) continue;
            const int side = (int)(c >> (nbits - 1 - bit)) & 1;  // first eligible axis = most significant bit
            ++bit;
            const IV sc = get_subinterval(IV(ce[2 * ax], ce[2 * ax + 1]), side, 2);
            ce[2 * ax] = sc.lb;
            ce[2 * ax + 1] = sc.ub;
            if (ax == 0) cla = cla * 2 + side;
          }
#pragma unroll
          for (int k = 0; k < 8; ++k) P.out.e(k, o) = ce[k];
          P.out.root(o) = root;
          P.out.la_idx(o) = cla;
          P.out.meta(o) = cmeta;
        });
      }
    }
  }
  // wave totals (wave-uniform) enter the block sum; PROVEN, EMPTY and OOD
  // items are certified, CRITICAL items that were not split, NOFIT and TRAP
  // items are unresolved
  uint64_t c_eval = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) c_eval += c_v[k];
  const uint64_t c_cert = (uint64_t)c_v[V_PROVEN] + c_v[V_EMPTY] + c_v[V_OOD];
  const uint64_t c_unres = (uint64_t)(c_v[V_CRITICAL] - c_split) + c_v[V_NOFIT] + c_v[V_TRAP];
  const unsigned long long acc[16] = {c_eval, c_v[0], c_v[1], c_v[2], c_v[3], c_v[4], c_v[5], 0,
                                      0,      c_dig,  c_vdig, c_cert, c_unres, c_f32, c_f32t, 0};
  wave_stats_add(stat_slot(P, P.depth, F, wave_slot()), acc);
}

// Root round with admissibility decided once per (la, r1, r2) triple: an
// out-of-domain triple accounts for its grid[3] root items (counts, digests,
// flags) without evaluating them; the items of a live triple are appended to
// `P.out` (depth 0) for k_round to evaluate.  The resulting counts, digests and
// flags equal those of evaluating every root item.
template <int F>
__global__ __launch_bounds__(1024) void k_root_triples(RefineParams P) {
  const int lane = threadIdx.x & 63;
  const uint64_t n_tri = (uint64_t)P.grid[0] * P.grid[1] * P.grid[2];
  const int n3 = P.grid[3];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const SegOut so{P.out_seg, P.seg_cap, P.overflow};
  SegCursor cur = seg_cursor();
  unsigned long long c_ood = 0, c_dig = 0, c_vdig = 0, c_live = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n_tri; base += stride) {
    uint64_t t = base + threadIdx.x;
    bool live = false;
    double e[8];
    int i0 = 0;
    if (t < n_tri) {
      uint64_t r = t;
      int i2 = (int)(r % (uint64_t)P.grid[2]); r /= (uint64_t)P.grid[2];
      int i1 = (int)(r % (uint64_t)P.grid[1]); r /= (uint64_t)P.grid[1];
      i0 = (int)r;
      e[0] = P.la_table[3 * i0 + 0];
      e[1] = P.la_table[3 * i0 + 1];
      IV r1 = get_subinterval(IV(P.box_lb[1], P.box_ub[1]), i1, P.grid[1]);
      IV r2 = get_subinterval(IV(P.box_lb[2], P.box_ub[2]), i2, P.grid[2]);
      e[2] = r1.lb; e[3] = r1.ub; e[4] = r2.lb; e[5] = r2.ub;
      IV la(e[0], e[1]);
      bool ood = false;
      IV r1c = r1, r2c = r2;
      if (P.clip) {
        r1c = r1.intersect(compute_r1_range(la, P.la_table[3 * i0 + 2]));
        ood = r1c.empty();
        if (!ood) {
          r2c = r2.intersect(r2_range<F>(la, r1c));
          ood = r2c.empty();
        }
      }
      if constexpr (is_3d(F)) {
        if (!ood) ood = r3_range<F>(la, r2c).empty();
      }
      live = !ood;
      if (live) c_live += (unsigned long long)n3;
      if (ood) {
        c_ood += (unsigned long long)n3;
        for (int k = 0; k < n3; ++k) {
          uint32_t root = (uint32_t)(t * n3 + k);
          if (P.digest) {
            IV p = get_subinterval(IV(P.box_lb[3], P.box_ub[3]), k, n3);
            e[6] = p.lb; e[7] = p.ub;
            uint64_t h = item_hash(e, root, F, 0u);
            c_dig += h;
            c_vdig += mix64(h ^ (uint64_t)vbyte(V_OOD, 0, 0));
          }
          if (P.root_flags) flag_root(P.root_flags, (uint64_t)root * P.nfam + P.fam_slot, CPL_RF_OOD);
        }
      }
    }
    const SegTicket tk = seg_reserve(so, cur, live, (uint32_t)n3);
    if (tk.total)
      seg_commit(so, cur, tk, [&](uint32_t k, uint64_t o) {
        const IV p = get_subinterval(IV(P.box_lb[3], P.box_ub[3]), (int)k, n3);
#pragma unroll
        for (int j = 0; j < 6; ++j) P.out.e(j, o) = e[j];
        P.out.e(6, o) = p.lb;
        P.out.e(7, o) = p.ub;
        P.out.root(o) = (uint32_t)(t * n3 + k);
        P.out.la_idx(o) = (uint32_t)i0;
        P.out.meta(o) = 0u;
      });
  }
  unsigned long long acc[16] = {c_ood, 0, 0, 0, 0, 0, 0, 0, 0, c_dig, c_vdig, c_ood, 0, 0, 0, c_live};
  acc[1 + V_OOD] = c_ood;
  block_stats_add(stat_slot(P, 0, F, blockIdx.x), acc);
}

// Zero family F's statistics records (every round and slot), the segment
// counts of the five worklist regions and the overflow word of family slot
// `slot`: the head of that family's captured chain in the graph form, so the
// chains need no common prologue (and no cross-stream wait) before they start.
__global__ void k_clear_family(RoundStats* stats, int F, unsigned long long* seg_counts, uint32_t* overflow, int slot) {
  constexpr int W = sizeof(RoundStats) / sizeof(unsigned long long);
  constexpr int per_round = STAT_SLOTS * W;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < CPL_MAX_ROUNDS * per_round) {
    const int d = i / per_round, k = i % per_round;
    reinterpret_cast<unsigned long long*>(stats + (size_t)(d * CPL_NUM_FAMILIES + F) * STAT_SLOTS)[k] = 0ull;
  }
  if (i < 5 * kSegs)
    seg_counts[((size_t)(i / kSegs) * CPL_NUM_FAMILIES + slot) * kSegs * kSegStride + (size_t)(i % kSegs) * kSegStride] = 0ull;
  if (i == 0) overflow[slot] = 0u;
}

// Item totals (sum of the clamped segment counts) of the regions of `nf`
// family slots whose counts start at `seg_counts`; one wave per slot.
__global__ void k_seg_totals(const unsigned long long* seg_counts, uint32_t seg_cap, uint64_t* totals) {
  uint64_t incl;
  const uint64_t n = seg_scan(seg_counts + (size_t)blockIdx.x * kSegs * kSegStride, seg_cap, incl);
  if (threadIdx.x == 0) totals[blockIdx.x] = n;
}

// Sum the STAT_SLOTS copies of every (round, family) record into out.
__global__ void k_reduce_stats(const RoundStats* __restrict__ in, RoundStats* __restrict__ out) {
  constexpr int W = sizeof(RoundStats) / sizeof(unsigned long long);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (record, word)
  if (i >= CPL_MAX_ROUNDS * CPL_NUM_FAMILIES * W) return;
  const int rec = i / W, w = i % W;
  const unsigned long long* src = reinterpret_cast<const unsigned long long*>(in + (size_t)rec * STAT_SLOTS) + w;
  unsigned long long acc = 0;
  for (int k = 0; k < STAT_SLOTS; ++k) acc += src[(size_t)k * W];
  reinterpret_cast<unsigned long long*>(out + rec)[w] = acc;
}

// Fill n words with a value (graph-safe replacement of hipMemsetAsync).
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ============================================================================
// Persistent form: one launch per family drains a mixed-depth queue.
//
// Each block is one wave that owns a depth-first stack of pending cells: the
// top lives in LDS (SoA), overflow goes to a private HBM region that only the
// same wave reads back.  A wave pops up to 64 cells (any mix of depths),
// evaluates them, pushes the children of critical cells (ballot + mbcnt ranks)
// and refills from the family's live root cells with a device-wide claim
// counter, in units that keep the batch full.  Waves never wait for each
// other: every iteration either evaluates cells, claims roots, or exits, so
// the kernel always terminates; capacity overflow is reported, not waited on.
// (The L2 caches of the XCDs are not coherent for ordinary device memory, so
// the only cross-wave communication is the atomic root counter.)
// ============================================================================
struct PersistParams {
  RefineParams R;
  Worklist roots;              // live root cells of this family (depth 0)
  const unsigned long long* root_seg;  // their segment counts
  uint32_t* q;                 // [0] root claim counter, [4] error flags
  double* sp_e;                // private spill stacks: [wave][8][spill_cap]
  uint32_t* sp_u;              // [wave][3][spill_cap]
  uint32_t spill_cap;          // cells per wave (multiple of 64)
  int cap;                     // cells in a wave's LDS stack (multiple of 64)
  uint32_t claim;              // roots claimed per atomic (small, so every wave gets a share)
};

enum : uint32_t { Q_ROOT = 0, Q_ERR = 4 };
enum : int {
  ST_EVAL = 0, ST_V = 1, ST_DIG = 9, ST_VDIG = 10, ST_CERT = 11, ST_UNRES = 12, ST_F32 = 13, ST_F32T = 14, ST_N = 15
};

// Child c of a cell under the refinement rule (same rule as k_round).
__device__ __forceinline__ void make_child(const RefineParams& P, const double e[8], uint32_t meta, uint32_t la_idx,
                                           int axis, int c, double ce[8], uint32_t& cmeta, uint32_t& cla) {
  const int depth = (int)(meta >> 16);
  const uint32_t splits = meta & 0xFFFFu;
#pragma unroll
  for (int k = 0; k < 8; ++k) ce[k] = e[k];
  cla = la_idx;
  if (P.split_mode == 0) {
    const int best = axis;
    IV sub = get_subinterval(IV(e[2 * best], e[2 * best + 1]), c, 2);
    ce[2 * best] = sub.lb;
    ce[2 * best + 1] = sub.ub;
    cmeta = ((uint32_t)(depth + 1) << 16) | (splits + (1u << (4 * best)));
    if (best == 0) cla = la_idx * 2 + c;
  } else {
    const int nbits = __popc(P.split_axes);
    uint32_t cs = splits;
    int bit = 0;
    for (int a = 0; a < 4; ++a) {
      if (!(P.split_axes & (1u << a))) continue;
      cs += (1u << (4 * a));
      int side = (c >> (nbits - 1 - bit)) & 1;
      ++bit;
      IV sub = get_subinterval(IV(ce[2 * a], ce[2 * a + 1]), side, 2);
      ce[2 * a] = sub.lb;
      ce[2 * a + 1] = sub.ub;
      if (a == 0) cla = cla * 2 + side;
    }
    cmeta = ((uint32_t)(depth + 1) << 16) | cs;
  }
}

// ---- training records of the learned proposal ----
struct RecordSink {
  float* feat;       // [cap][8]
  float* targ;       // [cap][8]
  uint32_t* count;
  uint64_t cap;
  uint32_t keep;     // keep one item in `keep` (by item hash)
};

template <bool ROOT>
__device__ __forceinline__ void load_item(const RefineParams& P, const Worklist& in, uint64_t i, double* e,
                                          uint32_t& root, uint32_t& la_idx, uint32_t& meta) {
  if (ROOT) {
    uint64_t r = i;
    int idx[4];
#pragma unroll
    for (int a = 3; a >= 0; --a) { idx[a] = (int)(r % (uint64_t)P.grid[a]); r /= (uint64_t)P.grid[a]; }
    e[0] = P.la_table[3 * idx[0] + 0];
    e[1] = P.la_table[3 * idx[0] + 1];
#pragma unroll
    for (int a = 1; a < 4; ++a) {
      IV s = get_subinterval(IV(P.box_lb[a], P.box_ub[a]), idx[a], P.grid[a]);
      e[2 * a] = s.lb;
      e[2 * a + 1] = s.ub;
    }
    root = (uint32_t)i;
    la_idx = (uint32_t)idx[0];
    meta = 0;
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = in.e(k, i);
    root = in.root(i);
    la_idx = in.la_idx(i);
    meta = in.meta(i);
  }
}

__device__ __forceinline__ double item_lbrsp(const RefineParams& P, uint32_t meta, uint32_t la_idx) {
  return P.clip ? P.la_table[3 * (la_level_offset(P.grid[0], axis_splits(meta, 0)) + la_idx) + 2] : 0.0;
}

// One record per kept item that reaches the criteria: features, whether the
// FP32 prefilter proves it, and (CRITICAL items that will be split) the
// settled fraction of the children of each eligible axis.
template <int F, bool ROOT>
__global__ __launch_bounds__(256) void k_record(RefineParams P, Worklist in, uint64_t n, RecordSink S) {
  uint64_t incl = 0;
  if (!ROOT && P.in_seg) n = seg_scan(P.in_seg, P.seg_cap, incl);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t a = (!ROOT && P.in_seg) ? seg_locate_run(base + (threadIdx.x & ~63u), i, incl, P.seg_cap) : i;
    if (i >= n) continue;
    double e[8];
    uint32_t root, la_idx, meta;
    load_item<ROOT>(P, in, a, e, root, la_idx, meta);
    if (S.keep > 1 && mix64(item_hash(e, root, F, meta)) % S.keep != 0) continue;
    const int depth = (int)(meta >> 16);
    const double lbrsp = item_lbrsp(P, meta, la_idx);
    const IV la(e[0], e[1]), r1(e[2], e[3]), r2(e[4], e[5]), p3(e[6], e[7]);
    const uint8_t vb = eval_synthetic<F>(la, r1, r2, p3, P.clip != 0, lbrsp);
    const int v = vb & 7;
    if (v != V_PROVEN && v != V_CRITICAL && v != V_TRAP) continue;
    bool hit = false;
    (void)eval_synthetic<F, true>(la, r1, r2, p3, P.clip != 0, lbrsp, true, &hit);
    float t[8] = {hit ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, (float)vb, (float)depth};
    if (v == V_CRITICAL && depth < P.max_depth && P.split_mode == 0) {
      t[5] = 1.0f;
      for (int a = 0; a < 4; ++a) {
        if (!(P.split_axes & (1u << a))) continue;
        int settled = 0;
        for (int c = 0; c < 2; ++c) {
          double ce[8];
          uint32_t cmeta, cla;
          make_child(P, e, meta, la_idx, a, c, ce, cmeta, cla);
          const int cv = eval_synthetic<F>(IV(ce[0], ce[1]), IV(ce[2], ce[3]), IV(ce[4], ce[5]), IV(ce[6], ce[7]),
                                           P.clip != 0, item_lbrsp(P, cmeta, cla)) & 7;
          settled += (cv == V_PROVEN || cv == V_EMPTY || cv == V_OOD) ? 1 : 0;
        }
        t[1 + a] = 0.5f * (float)settled;
      }
    }
    float f[8];
    mlp_features(e, mlp_box(P.box_lb, P.box_ub), f);
    const uint32_t pos = atomicAdd(S.count, 1u);
    if (pos < S.cap) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        S.feat[(size_t)pos * 8 + k] = f[k];
        S.targ[(size_t)pos * 8 + k] = t[k];
      }
    }
  }
}

template <int F, bool EXT = false>
__global__ __launch_bounds__(64) void k_persist(PersistParams Q) {
  extern __shared__ __align__(16) char lds[];
  const RefineParams& P = Q.R;
  const int CAP = Q.cap;
  const int lane = threadIdx.x;
  double* se = reinterpret_cast<double*>(lds);                        // [8][CAP]
  uint32_t* su = reinterpret_cast<uint32_t*>(se + 8 * CAP);           // [3][CAP]
  unsigned long long* sst = reinterpret_cast<unsigned long long*>(su + 3 * CAP);  // [MAX_ROUNDS][ST_N]
  for (int k = lane; k < CPL_MAX_ROUNDS * ST_N; k += 64) sst[k] = 0;
  const uint32_t SC = Q.spill_cap;
  double* pe = Q.sp_e + (size_t)blockIdx.x * 8 * SC;
  uint32_t* pu = Q.sp_u + (size_t)blockIdx.x * 3 * SC;
  const int NCH = P.n_children;
  const int S = 64 * NCH;        // most cells one batch can push
  uint64_t root_incl;
  const uint32_t n_roots = (uint32_t)seg_scan(Q.root_seg, P.seg_cap, root_incl);  // overflow is reported by the host
  int count = 0;                 // cells in LDS (wave-uniform)
  uint32_t spilled = 0;          // cells in the private HBM stack (wave-uniform, multiple of 64)
  bool roots_left = true;

  for (;;) {
    // ---- refill: own spilled cells first (depth first), then new roots ----
    if (count < 64 && spilled > 0) {
      spilled -= 64;
      for (int k = 0; k < 8; ++k) se[k * CAP + count + lane] = pe[(size_t)k * SC + spilled + lane];
      for (int k = 0; k < 3; ++k) su[k * CAP + count + lane] = pu[(size_t)k * SC + spilled + lane];
      count += 64;
    }
    while (count < 64 && roots_left) {
      const uint32_t want = min((uint32_t)(64 - count), Q.claim);
      uint32_t r0 = 0;
      if (lane == 0) r0 = atomicAdd(Q.q + Q_ROOT, want);
      r0 = __shfl(r0, 0, 64);
      if (r0 >= n_roots) { roots_left = false; break; }
      const uint32_t take = min(want, n_roots - r0);
      const uint64_t i = seg_locate(r0 + min((uint32_t)lane, take - 1), root_incl, P.seg_cap);
      if ((uint32_t)lane < take) {
        for (int k = 0; k < 8; ++k) se[k * CAP + count + lane] = Q.roots.e(k, i);
        su[0 * CAP + count + lane] = Q.roots.root(i);
        su[1 * CAP + count + lane] = Q.roots.la_idx(i);
        su[2 * CAP + count + lane] = Q.roots.meta(i);
      }
      count += (int)take;
    }
    if (count == 0) break;  // nothing in LDS, nothing spilled, no roots left

    // ---- evaluate a batch from the top of the stack ----
    const int nb = count < 64 ? count : 64;
    const int base = count - nb;
    const bool valid = lane < nb;
    double e[8];
    uint32_t root = 0, la_idx = 0, meta = 0;
    if (valid) {
      for (int k = 0; k < 8; ++k) e[k] = se[k * CAP + base + lane];
      root = su[0 * CAP + base + lane];
      la_idx = su[1 * CAP + base + lane];
      meta = su[2 * CAP + base + lane];
    }
    count = base;
    bool split = false;
    const uint8_t pb = kNoProposal;  // the learned proposal runs with the round-based forms only
    int depth = 0, v = 0, cert = 0;
    bool hit = false, tried = false;
    uint64_t h = 0, hv = 0;
    if (valid) {
      depth = (int)(meta >> 16);
      double lbrsp = 0.0;
      if (P.clip) lbrsp = P.la_table[3 * (la_level_offset(P.grid[0], axis_splits(meta, 0)) + la_idx) + 2];
      const bool pre = EXT && ((P.prefilter_depths >> depth) & 1u) && (pb & 1);
      uint8_t vb = eval_synthetic<F, EXT>(IV(e[0], e[1]), IV(e[2], e[3]), IV(e[4], e[5]), IV(e[6], e[7]), P.clip != 0,
                                          lbrsp, pre, &hit, &tried);
      v = vb & 7;
      if (P.digest) {
        h = item_hash(e, root, F, meta);
        hv = mix64(h ^ vb);
      }
      uint32_t fl = (vb & 8) ? CPL_RF_NOFIT_FLAG : 0;
      if (v == V_CRITICAL && depth < P.max_depth) {
        split = true;
      } else {
        switch (v) {
          case V_PROVEN: fl |= CPL_RF_PROVEN; cert = 1; break;
          case V_EMPTY: fl |= CPL_RF_EMPTY; cert = 1; break;
          case V_OOD: fl |= CPL_RF_OOD; cert = 1; break;
          case V_CRITICAL: fl |= CPL_RF_CRITICAL; cert = 2; break;
          case V_NOFIT: fl |= CPL_RF_NOFIT; cert = 2; break;
          default: fl |= CPL_RF_TRAP; cert = 2; break;
        }
      }
      if (fl && P.root_flags) flag_root(P.root_flags, (uint64_t)root * P.nfam + P.fam_slot, fl);
    }
    // ---- statistics: per depth present in the batch, counts from ballots, sums
    // by shuffles; the block is one wave, so its LDS record needs no atomics ----
    for (uint64_t rem = __ballot(valid); rem;) {
      const int d0 = __shfl(depth, __ffsll((unsigned long long)rem) - 1, 64);
      const bool in = valid && depth == d0;
      const uint64_t m = __ballot(in);
      rem &= ~m;
      unsigned long long cnt[ST_N];
      cnt[ST_EVAL] = __popcll(m);
#pragma unroll
      for (int k = 0; k < 8; ++k) cnt[ST_V + k] = __popcll(__ballot(in && v == k));
      cnt[ST_CERT] = __popcll(__ballot(in && cert == 1));
      cnt[ST_UNRES] = __popcll(__ballot(in && cert == 2));
      cnt[ST_F32] = __popcll(__ballot(in && hit));
      cnt[ST_F32T] = __popcll(__ballot(in && tried));
      unsigned long long dg = in ? h : 0, dv = in ? hv : 0;
      if (P.digest) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          dg += __shfl_xor(dg, off, 64);
          dv += __shfl_xor(dv, off, 64);
        }
      }
      cnt[ST_DIG] = dg;
      cnt[ST_VDIG] = dv;
      if (lane == 0) {
        unsigned long long* st = sst + d0 * ST_N;
#pragma unroll
        for (int k = 0; k < ST_N; ++k) st[k] += cnt[k];
      }
    }
    // ---- push children ----
    const uint64_t mask = __ballot(split);
    if (mask) {
      const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
      if (split) {
        const int axis = P.split_mode == 0 ? split_axis(P, meta, pb) : 0;
        for (int c = 0; c < NCH; ++c) {
          double ce[8];
          uint32_t cmeta, cla;
          make_child(P, e, meta, la_idx, axis, c, ce, cmeta, cla);
          const int o = count + rank * NCH + c;
          for (int k = 0; k < 8; ++k) se[k * CAP + o] = ce[k];
          su[0 * CAP + o] = root;
          su[1 * CAP + o] = cla;
          su[2 * CAP + o] = cmeta;
        }
      }
      count += (int)__popcll(mask) * NCH;
    }
    // ---- spill the top 64 cells to the private stack while LDS could overflow ----
    bool full = false;
    while (count > CAP - S) {
      if (spilled + 64 > SC) { full = true; break; }
      const int from = count - 64;
      for (int j = 0; j < 8; ++j) pe[(size_t)j * SC + spilled + lane] = se[j * CAP + from + lane];
      for (int j = 0; j < 3; ++j) pu[(size_t)j * SC + spilled + lane] = su[j * CAP + from + lane];
      spilled += 64;
      count -= 64;
    }
    if (full) {
      if (lane == 0) atomicOr(Q.q + Q_ERR, 1u);
      break;
    }
  }
  // flush the wave's statistics
  __syncthreads();
  for (int k = lane; k < CPL_MAX_ROUNDS * ST_N; k += 64) {
    unsigned long long v = sst[k];
    if (!v) continue;
    const int d = k / ST_N, c = k % ST_N;
    RoundStats* S = stat_slot(P, d, F, blockIdx.x);
    unsigned long long* dst = c == ST_EVAL ? &S->evaluations
                            : c < ST_DIG   ? &S->verdicts[c - ST_V]
                            : c == ST_DIG  ? &S->digest
                            : c == ST_VDIG ? &S->vdigest
                            : c == ST_CERT ? &S->certified
                            : c == ST_F32  ? &S->fp32_settled
                            : c == ST_F32T ? &S->fp32_tried
                                           : &S->unresolved;
    atomicAdd(dst, v);
  }
}

template <int F>
__global__ __launch_bounds__(256) void k_eval_cells(int clip, const double* e0, const double* e1, const double* e2,
                                                     const double* e3, const double* e4, const double* e5,
                                                     const double* e6, const double* e7, size_t n, uint8_t* out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  IV la(e0[i], e1[i]);
  double lbrsp = clip ? lower_bound_radius_single_placement(la.lb) : 0.0;
  out[i] = eval_synthetic<F>(la, IV(e2[i], e3[i]), IV(e4[i], e5[i]), IV(e6[i], e7[i]), clip != 0, lbrsp);
}

template <int F>
__global__ __launch_bounds__(256) void k_body(const double* cells, size_t n, uint8_t* out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* c = cells + 8 * i;
  out[i] = eval_kernel_body<F>(IV(c[0], c[1]), IV(c[2], c[3]), IV(c[4], c[5]), IV(c[6], c[7]));
}

// ---- reference grid (Mode R) ----------------------------------------------
// Scratch of one launch (device memory, MODE_R_MISC_BYTES): the ordered-u64
// minimum of failing la.lb (Atomic_ordered_double), the smallest la.lb of each
// failure kind, the record counter, and 64 statistics slots of one 128-byte
// line each (per wave: cells evaluated, rows skipped, start no-fit flags).
enum : int { MR_FIRST = 0, MR_FAIL = 4, MR_NREC = 10, MR_SLOTS = 32, MR_SLOT_STRIDE = 16, MR_NSLOTS = 64 };
constexpr size_t MODE_R_MISC_BYTES = (MR_SLOTS + MR_NSLOTS * MR_SLOT_STRIDE) * 8;

struct ModeRParams {
  IV outer;
  int n[4];                   // subdivisions la, r1, r2, r3 (r3 = 1 for 2-d families)
  const double* r1_ranges;    // per la index: la.lb, la.ub, r1_range.lb, r1_range.ub
  unsigned long long* misc;   // MODE_R_MISC_BYTES, laid out as MR_*
  int exhaustive;             // evaluate every cell (no early exit)
  double* crit_out;           // output mode: records of 9 doubles (kind, la, r1, r2, r3)
  uint32_t crit_cap;
};

__global__ void k_mode_r_ranges(IV outer, int nla, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nla) return;
  IV la = get_subinterval(outer, i, nla);
  IV r = compute_r1_range(la, lower_bound_radius_single_placement(la.lb));
  out[4 * i + 0] = la.lb;
  out[4 * i + 1] = la.ub;
  out[4 * i + 2] = r.lb;
  out[4 * i + 3] = r.ub;
}

// Clears a launch's scratch (a kernel, not a memset, so the call allocates and
// records nothing but kernels): first = DBL_MAX (run_kernel.cuh:81), failure
// minima = none, counters = 0.
__global__ void k_mode_r_init(unsigned long long* misc) {
  for (int i = threadIdx.x; i < (int)(MODE_R_MISC_BYTES / 8); i += blockDim.x) {
    unsigned long long v = 0;
    if (i == MR_FIRST) v = ordered_of(DBL_MAX);
    else if (i >= MR_FAIL && i < MR_FAIL + 6) v = ~0ull;
    misc[i] = v;
  }
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One launch of the reference kernel of family F over the reference grid, in
// the reference's cell order (r3 fastest, then r2, r1, la).  Each wave takes
// (la, r1) rows in turn (grid stride over rows by the global wave index, so
// the whole grid sweeps la upwards): it computes the row's r1 subinterval, r2
// range and semicircle angle of r1 once, and its lanes walk the row's n2 * n3
// cells 64 apart with a mixed-radix (i2, i3) counter.
//
// Without output (the reference's store_min + return, start_configuration.cu:
// 142-145 and every kernel alike): a reference thread walks its cells in
// increasing la and returns at its first failing cell, so cells whose la.lb is
// at or above a failing la.lb already found cannot change the minimum.  A wave
// skips such rows (the global minimum is re-read every row and every 16 cells
// of a long row with a relaxed load; a stale value only skips less) and stops
// a row once one of its lanes failed in it.  Failures other than a failing
// criterion (trap codes, the middle and end families' fit check) are recorded
// with their la.lb; the host reports those below the final minimum, the cells
// every reference thread is bound to reach.  P.exhaustive turns the skipping
// off (every cell evaluated, as the counts of parity tests need).
//
// OUT (output_criticals): every cell and every criterion is evaluated and one
// record is written per printed line (kind 1/2: failing criterion, 3: the start
// family's no-fit line); every failure counts.
template <int F, bool OUT>
__global__ __launch_bounds__(256, 4) void k_mode_r(ModeRParams P) {
  const uint32_t n1 = (uint32_t)P.n[1], n2 = (uint32_t)P.n[2], n3 = (uint32_t)P.n[3];
  const uint64_t rows = (uint64_t)P.n[0] * n1;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  // start cell and step of this lane within a row, as (i2, i3) digits
  const uint32_t s2 = lane / n3, s3 = lane % n3;
  const uint32_t d2 = 64 / n3, d3 = 64 % n3;
  unsigned long long* first = P.misc + MR_FIRST;
  unsigned long long best = ~0ull;       // this lane's smallest failing ordered la.lb
  unsigned long long published = ~0ull;  // the wave's smallest value published to *first
  unsigned long long ev = 0, skipped = 0, nofit_flags = 0;
  for (uint64_t row = wave; row < rows; row += waves) {
    const int ia = (int)(row / n1), i1 = (int)(row % n1);
    const double* rr = P.r1_ranges + 4 * ia;
    const IV la(rr[0], rr[1]);
    const unsigned long long key = ordered_of(la.lb);
    if (!OUT && !P.exhaustive) {
      const unsigned long long g = __hip_atomic_load(first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__ballot(key >= (g < published ? g : published))) {
        skipped = (rows - 1 - row) / waves + 1;  // this row and the later ones of the wave: la only grows
        break;
      }
    }
    const IV r1_ = get_subinterval(IV(rr[2], rr[3]), i1, P.n[1]);
    const IV r2rng = r2_range<F>(la, r1_);
    // the row's semicircle angle of r1 for the criteria that use it on most cells
    // (the start family's criteria reach it rarely; holding it there costs more)
    constexpr bool kA1 = !OUT && (F == F_START_END_2D || F == F_START_END_3D || F == F_END_2D);
    IV a1r(0.0, 0.0);
    if constexpr (kA1) a1r = semicircle_angle_outer(r1_);
    uint32_t iter = 0;
    for (uint32_t i2 = s2, i3 = s3; __ballot(i2 < n2); ++iter) {
      if (!OUT && !P.exhaustive) {
        // the rest of the row cannot lower the minimum once a lane failed in it
        if (__ballot(best <= key)) break;
        if (iter && (iter & 15) == 0) {
          const unsigned long long g = __hip_atomic_load(first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (__ballot(key >= g)) break;
        }
      }
      const bool active = i2 < n2;
      const uint32_t c2 = i2, c3 = i3;
      // next cell of this lane (i3 + d3 < 2 n3, one carry at most)
      i3 += d3;
      i2 += d2;
      if (i3 >= n3) { i3 -= n3; ++i2; }
      if (!active) continue;
      const IV r2_ = get_subinterval(r2rng, (int)c2, P.n[2]);
      IV r3(0.0, 0.0);
      if constexpr (is_3d(F)) {
        IV range = r3_range<F>(la, r2_);
        if (F == F_END_3D && range.empty()) continue;  // end_configuration.cu:222-224
        r3 = get_subinterval(range, (int)c3, P.n[3]);
      }
      ++ev;
      if constexpr (!OUT) {
        const uint8_t vb = eval_kernel_body<F>(la, r1_, r2_, r3, true, nullptr, nullptr, kA1 ? &a1r : nullptr);
        const int v = vb & 7;
        nofit_flags += (vb >> 3) & 1;
        if (v == V_CRITICAL) {
          best = key < best ? key : best;
        } else if (v == V_NOFIT || v == V_TRAP) {
          atomicMin(P.misc + MR_FAIL + (v == V_NOFIT ? 0 : min((int)(vb >> 4), 5)), key);
        }
      } else {
        int nofit = 0, early = -1;
        const int res = eval_kernel_body_all<F>(la, r1_, r2_, r3, &nofit, &early);
        if (early == V_NOFIT) atomicMin(P.misc + MR_FAIL, key);
        if (early >= 0) continue;
        if (res >= 16) atomicMin(P.misc + MR_FAIL + min(res - 16, 5), key);
        const int mask = res >= 16 ? 0 : res;
        if (mask) best = key < best ? key : best;
        nofit_flags += nofit;
        if (nofit | mask) {
          IV r2 = r2_;
          if (is_3d(F)) r2.tighten_lb(r3.lb);
          IV r1 = r1_;
          r1.tighten_lb(r2.lb);
          for (int kind = 1; kind <= 3; ++kind) {
            if (!(kind == 3 ? nofit : (mask >> (kind - 1)) & 1)) continue;
            const unsigned long long slot = atomicAdd(P.misc + MR_NREC, 1ull);
            if (slot < P.crit_cap) {
              double* rec = P.crit_out + 9 * (size_t)slot;
              rec[0] = (double)kind;
              rec[1] = la.lb; rec[2] = la.ub; rec[3] = r1.lb; rec[4] = r1.ub;
              rec[5] = r2.lb; rec[6] = r2.ub; rec[7] = r3.lb; rec[8] = r3.ub;
            }
          }
        }
      }
    }
    if (__ballot(best < published)) {  // publish the wave's new minimum at once
      published = wave_min_u64(best);
      if (lane == 0) atomicMin(first, published);
    }
  }
  ev = wave_sum_u64(ev);
  nofit_flags = wave_sum_u64(nofit_flags);
  if (lane == 0) {
    unsigned long long* slot = P.misc + MR_SLOTS + MR_SLOT_STRIDE * (wave & (MR_NSLOTS - 1));
    if (ev) atomicAdd(slot + 0, ev);
    if (skipped) atomicAdd(slot + 1, skipped);
    if (nofit_flags) atomicAdd(slot + 2, nofit_flags);
  }
}

// Device scratch of cpl_first_problematic_noalloc: the per-lambda r1 ranges
// of up to CPL_NOALLOC_MAX_LA lambda subdivisions, one launch's MODE_R_MISC_BYTES
// and up to CPL_NOALLOC_MAX_RECORDS output-mode records.  Every GPU the library
// runs on holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns at initialisation.
constexpr size_t kNoAllocRangesBytes = (size_t)CPL_NOALLOC_MAX_LA * 4 * sizeof(double);
constexpr size_t kNoAllocRecordBytes = (size_t)CPL_NOALLOC_MAX_RECORDS * 9 * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocRangesBytes + MODE_R_MISC_BYTES + kNoAllocRecordBytes;
static_assert(kNoAllocRangesBytes % 256 == 0 && MODE_R_MISC_BYTES % 256 == 0, "scratch sections stay aligned");
static_assert(kNoAllocBytes == CPL_NOALLOC_SCRATCH_BYTES, "the scratch size cpl.h states");
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// ============================================================================
// Host side: context and C ABI
// ============================================================================
namespace {

using cpl_host::DevBuf;
using cpl_host::fail;
using cpl_host::g_err;

constexpr size_t kStageStats = (size_t)CPL_MAX_ROUNDS * CPL_NUM_FAMILIES * sizeof(RoundStats);
constexpr size_t kStageBytes = kStageStats + CPL_NUM_FAMILIES * sizeof(uint32_t);

struct Context {
  bool ready = false;
  int device = 0;
  int cu_count = 0;
  unsigned char* noalloc_scratch = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress)
  hipStream_t stream = nullptr;
  hipStream_t fam_stream[CPL_NUM_FAMILIES] = {};   // per-family branches (default priority)
  hipEvent_t fork_ev = nullptr, join_ev[CPL_NUM_FAMILIES] = {};
  hipEvent_t chain_end[CPL_NUM_FAMILIES] = {};  // graph form: end of each family chain (timed)
  hipEvent_t run_begin = nullptr, run_end = nullptr;  // graph form: the replay's bounds (timed)
  // pinned host copy of the summed statistics records and the families'
  // overflow words, read back with one copy after a graph replay (allocated
  // by the first refinement run)
  unsigned char* stage = nullptr;
  DevBuf scratch_in, scratch_in2, scratch_out, cells, verdicts;
  DevBuf wl[2];           // ping-pong worklists; each holds nfam regions
  DevBuf la_table, root_flags, stats, stats_sum, counters;
  DevBuf mr_ranges, mr_misc, mr_crit;
  DevBuf persist_q;
  DevBuf mlp;                 // MlpPacked per family
  DevBuf prop;                // proposal bytes of a round's input, one region per family
  bool mlp_ready[CPL_NUM_FAMILIES] = {};
  DevBuf records;             // training records being collected
  bool recording = false;
  RecordSink sink{};
  uint64_t wl_cap_items = 0;  // per family region
};
Context g;

int ensure_init() {
  if (g.ready) return CPL_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(CPL_ERR_NO_DEVICE, "no HIP device available");
  HIP_TRY(hipGetDevice(&g.device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, g.device));
  g.cu_count = prop.multiProcessorCount;
  HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  for (int f = 0; f < CPL_NUM_FAMILIES; ++f) {
    HIP_TRY(hipStreamCreateWithFlags(&g.fam_stream[f], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&g.join_ev[f], hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&g.chain_end[f]));
  }
  HIP_TRY(hipEventCreateWithFlags(&g.fork_ev, hipEventDisableTiming));
  HIP_TRY(hipEventCreate(&g.run_begin));
  HIP_TRY(hipEventCreate(&g.run_end));
  // the device address of this device's copy of the no-allocation scratch
  void* sp = nullptr;
  HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  if (!sp) return fail(CPL_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
  g.noalloc_scratch = static_cast<unsigned char*>(sp);
  g.ready = true;
  return CPL_OK;
}

}  // namespace

int cpl_host::ensure_init() { return ::ensure_init(); }
hipStream_t cpl_host::stream() { return g.stream; }

namespace {

// Stream the refinement launch helpers enqueue on (the library stream, or a
// family branch while forking).
hipStream_t g_launch_stream = nullptr;
hipStream_t launch_stream() { return g_launch_stream ? g_launch_stream : g.stream; }

unsigned grid_for(uint64_t n, int block) {
  uint64_t want = (n + block - 1) / block;
  uint64_t cap = (uint64_t)std::max(g.cu_count, 1) * 16;
  if (want < 1) want = 1;
  return (unsigned)std::min(want, cap);
}

template <template <int> class K, class... Args>
void dispatch_family(int f, Args&&... args) {
  switch (f) {
    case 0: K<0>::run(args...); break;
    case 1: K<1>::run(args...); break;
    case 2: K<2>::run(args...); break;
    case 3: K<3>::run(args...); break;
    case 4: K<4>::run(args...); break;
    default: K<5>::run(args...); break;
  }
}

template <int F> struct LaunchEvalCells {
  static void run(unsigned grid, int clip, double** d, size_t n, uint8_t* out) {
    hipLaunchKernelGGL(k_eval_cells<F>, dim3(grid), dim3(256), 0, g.stream, clip, d[0], d[1], d[2], d[3], d[4], d[5],
                       d[6], d[7], n, out);
  }
};
template <int F> struct LaunchBody {
  static void run(unsigned grid, const double* cells, size_t n, uint8_t* out) {
    hipLaunchKernelGGL(k_body<F>, dim3(grid), dim3(256), 0, g.stream, cells, n, out);
  }
};
template <int F> struct LaunchRoundB {
  template <bool R, int LB, bool EXT, int WAVES = 3>
  static void go(unsigned grid, int block, const RefineParams& P, const Worklist& in, uint64_t n) {
    hipLaunchKernelGGL((k_round<F, R, LB, EXT, WAVES>), dim3(grid), dim3(block), 0, launch_stream(), P, in, n);
  }
  static void run(bool root, bool p32, unsigned grid, int block, const RefineParams& P, const Worklist& in,
                  uint64_t n) {
    const int k = (root ? 4 : 0) | (block > 256 ? 2 : 0) | (p32 ? 1 : 0);
    switch (k) {
      case 0: go<false, 256, false>(grid, block, P, in, n); break;
      case 1: go<false, 256, true>(grid, block, P, in, n); break;
      case 2: go<false, 1024, false>(grid, block, P, in, n); break;
      case 3: go<false, 1024, true>(grid, block, P, in, n); break;
      case 4: go<true, 256, false>(grid, block, P, in, n); break;
      case 5: go<true, 256, true>(grid, block, P, in, n); break;
      case 6: go<true, 1024, false>(grid, block, P, in, n); break;
      default: go<true, 1024, true>(grid, block, P, in, n); break;
    }
  }
};
template <int F> struct LaunchRound {
  static void run(bool root, bool p32, unsigned grid, const RefineParams& P, const Worklist& in, uint64_t n) {
    LaunchRoundB<F>::run(root, p32, grid, 256, P, in, n);
  }
};
template <int F> struct LaunchPrefilterDebug {
  static void run(unsigned grid, const double* cells, size_t n, uint8_t* o32, uint8_t* o64) {
    hipLaunchKernelGGL(k_debug_prefilter<F>, dim3(grid), dim3(256), 0, g.stream, cells, n, o32, o64);
  }
};
template <int F> struct LaunchRecord {
  static void run(bool root, unsigned grid, const RefineParams& P, const Worklist& in, uint64_t n, const RecordSink& S) {
    if (root) hipLaunchKernelGGL((k_record<F, true>), dim3(grid), dim3(256), 0, g.stream, P, in, n, S);
    else hipLaunchKernelGGL((k_record<F, false>), dim3(grid), dim3(256), 0, g.stream, P, in, n, S);
  }
};
template <int F> struct LaunchModeR {
  static void run(bool out, unsigned grid, const ModeRParams& P) {
    if (out) hipLaunchKernelGGL((k_mode_r<F, true>), dim3(grid), dim3(256), 0, g.stream, P);
    else hipLaunchKernelGGL((k_mode_r<F, false>), dim3(grid), dim3(256), 0, g.stream, P);
  }
};

Worklist region(DevBuf& b, uint64_t cap, int slot) {
  // layout per region: 8 double arrays, then 3 u32 arrays
  const size_t per_item = 8 * sizeof(double) + 3 * sizeof(uint32_t);
  char* base = b.as<char>() + (size_t)slot * cap * per_item;
  Worklist w;
  w.ev = reinterpret_cast<double*>(base);
  w.uv = reinterpret_cast<uint32_t*>(base + 8 * cap * sizeof(double));
  w.cap = cap;
  return w;
}

}  // namespace

extern "C" {

const char* cpl_last_error(void) { return g_err.c_str(); }

int cpl_init(void) { return ensure_init(); }

namespace {
void drop_graphs();
}

int cpl_shutdown(void) {
  if (!g.ready) return CPL_OK;
  (void)hipStreamSynchronize(g.stream);
  drop_graphs();  // their kernels address the buffers released below
  for (bool& r : g.mlp_ready) r = false;  // the proposals live in g.mlp
  for (DevBuf* b : {&g.scratch_in, &g.scratch_in2, &g.scratch_out, &g.cells, &g.verdicts, &g.wl[0], &g.wl[1],
                    &g.la_table, &g.root_flags, &g.stats, &g.stats_sum, &g.counters, &g.mr_ranges, &g.mr_misc, &g.mr_crit,
                    &g.persist_q, &g.mlp, &g.records, &g.prop})
    b->release();
  (void)hipStreamDestroy(g.stream);
  g.stream = nullptr;
  for (int f = 0; f < CPL_NUM_FAMILIES; ++f) {
    (void)hipStreamDestroy(g.fam_stream[f]);
    (void)hipEventDestroy(g.join_ev[f]);
    (void)hipEventDestroy(g.chain_end[f]);
    g.fam_stream[f] = nullptr;
  }
  (void)hipEventDestroy(g.fork_ev);
  (void)hipEventDestroy(g.run_begin);
  (void)hipEventDestroy(g.run_end);
  (void)hipHostFree(g.stage);
  g.stage = nullptr;
  g.ready = false;
  return CPL_OK;
}

// FP64 FMA throughput probe (checks the published FP64 peak the roofline uses):
// eight independent fma chains per lane cover the pipeline latency; operands
// come from arguments and the chains converge to 1, so nothing folds and the
// final store never happens.
__global__ __launch_bounds__(256) void k_fp64_peak(int iters, double b, double c, double* out) {
  double a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = (double)(threadIdx.x + k) * 1e-3;
#pragma unroll 4
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = __builtin_fma(a[k], b, c);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  if (s == 12345.0) out[0] = s;
}

int cpl_debug_fp64_peak(int iters, double* tflops) {
  int rc = ensure_init();
  if (rc) return rc;
  if (iters <= 0 || !tflops) return fail(CPL_ERR_ARG, "cpl_debug_fp64_peak: bad argument");
  if ((rc = g.scratch_out.ensure(sizeof(double)))) return rc;
  const unsigned blocks = (unsigned)std::max(g.cu_count, 1) * 8;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_fp64_peak, dim3(blocks), dim3(256), 0, g.stream, std::min(iters, 1024), 0.999, 1e-3,
                     g.scratch_out.as<double>());  // warm-up
  HIP_TRY(hipEventRecord(e0, g.stream));
  hipLaunchKernelGGL(k_fp64_peak, dim3(blocks), dim3(256), 0, g.stream, iters, 0.999, 1e-3, g.scratch_out.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, g.stream));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double flops = (double)blocks * 256.0 * (double)iters * 8.0 * 2.0;
  *tflops = ms > 0.f ? flops / (ms * 1e-3) / 1e12 : 0.0;
  return CPL_OK;
}

int cpl_debug_prim(int op, const double* a, const double* b, double* out, size_t n) {
  int rc = ensure_init();
  if (rc) return rc;
  if (op < 0 || op > 41 || !a || !out) return fail(CPL_ERR_ARG, "cpl_debug_prim: bad op or pointer");
  if (n == 0) return CPL_OK;
  size_t bytes = n * sizeof(double);
  if ((rc = g.scratch_in.ensure(bytes)) || (rc = g.scratch_in2.ensure(bytes)) || (rc = g.scratch_out.ensure(bytes)))
    return rc;
  HIP_TRY(hipMemcpyAsync(g.scratch_in.p, a, bytes, hipMemcpyHostToDevice, g.stream));
  if (b) HIP_TRY(hipMemcpyAsync(g.scratch_in2.p, b, bytes, hipMemcpyHostToDevice, g.stream));
  hipLaunchKernelGGL(k_debug_prim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g.stream, op,
                     g.scratch_in.as<double>(), b ? g.scratch_in2.as<double>() : nullptr, g.scratch_out.as<double>(), n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, g.scratch_out.p, bytes, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  return CPL_OK;
}

int cpl_eval_cells(int family, int clip, const double* const soa[8], size_t n, uint8_t* verdicts) {
  int rc = ensure_init();
  if (rc) return rc;
  if (family < 0 || family >= CPL_NUM_FAMILIES || !soa || !verdicts) return fail(CPL_ERR_ARG, "cpl_eval_cells: bad argument");
  if (n == 0) return CPL_OK;
  const size_t bytes = n * sizeof(double);
  if ((rc = g.cells.ensure(8 * bytes)) || (rc = g.verdicts.ensure(n))) return rc;
  double* d[8];
  for (int k = 0; k < 8; ++k) {
    d[k] = g.cells.as<double>() + k * n;
    HIP_TRY(hipMemcpyAsync(d[k], soa[k], bytes, hipMemcpyHostToDevice, g.stream));
  }
  dispatch_family<LaunchEvalCells>(family, (unsigned)((n + 255) / 256), clip, d, n, g.verdicts.as<uint8_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(verdicts, g.verdicts.p, n, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  return CPL_OK;
}

int cpl_eval_kernel_body(int family, const double* cells, size_t n, uint8_t* verdicts) {
  int rc = ensure_init();
  if (rc) return rc;
  if (family < 0 || family >= CPL_NUM_FAMILIES || !cells || !verdicts) return fail(CPL_ERR_ARG, "cpl_eval_kernel_body: bad argument");
  if (n == 0) return CPL_OK;
  if ((rc = g.cells.ensure(8 * n * sizeof(double))) || (rc = g.verdicts.ensure(n))) return rc;
  HIP_TRY(hipMemcpyAsync(g.cells.p, cells, 8 * n * sizeof(double), hipMemcpyHostToDevice, g.stream));
  dispatch_family<LaunchBody>(family, (unsigned)((n + 255) / 256), g.cells.as<double>(), n, g.verdicts.as<uint8_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(verdicts, g.verdicts.p, n, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  return CPL_OK;
}

int cpl_debug_geom(int fn, int fp32, const double* in, double* out, int* trap, size_t n) {
  int rc = ensure_init();
  if (rc) return rc;
  if (fn < 0 || fn > 8 || !in || !out || !trap) return fail(CPL_ERR_ARG, "cpl_debug_geom: bad function or pointer");
  if (n == 0) return CPL_OK;
  if ((rc = g.scratch_in.ensure(6 * n * sizeof(double))) || (rc = g.scratch_out.ensure(2 * n * sizeof(double))) ||
      (rc = g.scratch_in2.ensure(n * sizeof(int))))
    return rc;
  HIP_TRY(hipMemcpyAsync(g.scratch_in.p, in, 6 * n * sizeof(double), hipMemcpyHostToDevice, g.stream));
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (fp32)
    hipLaunchKernelGGL(k_debug_geom<float>, dim3(grid), dim3(256), 0, g.stream, fn, g.scratch_in.as<double>(),
                       g.scratch_out.as<double>(), g.scratch_in2.as<int>(), n);
  else
    hipLaunchKernelGGL(k_debug_geom<double>, dim3(grid), dim3(256), 0, g.stream, fn, g.scratch_in.as<double>(),
                       g.scratch_out.as<double>(), g.scratch_in2.as<int>(), n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, g.scratch_out.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipMemcpyAsync(trap, g.scratch_in2.p, n * sizeof(int), hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  return CPL_OK;
}

int cpl_interval_op(int op, const cpl_iv* a, const cpl_iv* b, cpl_iv* out, int* trap, size_t n) {
  int rc = ensure_init();
  if (rc) return rc;
  if (op < 0 || op > 15 || !a || !out || !trap) return fail(CPL_ERR_ARG, "cpl_interval_op: bad op or pointer");
  if (op <= 3 || op == 8 || op == 9 || op == 14)
    if (!b) return fail(CPL_ERR_ARG, "cpl_interval_op: binary operation needs b");
  if (n == 0) return CPL_OK;
  const size_t bytes = n * sizeof(cpl_iv);
  if ((rc = g.scratch_in.ensure(bytes)) || (rc = g.scratch_in2.ensure(bytes)) || (rc = g.scratch_out.ensure(bytes)) ||
      (rc = g.verdicts.ensure(n * sizeof(int))))
    return rc;
  HIP_TRY(hipMemcpyAsync(g.scratch_in.p, a, bytes, hipMemcpyHostToDevice, g.stream));
  if (b) HIP_TRY(hipMemcpyAsync(g.scratch_in2.p, b, bytes, hipMemcpyHostToDevice, g.stream));
  hipLaunchKernelGGL(k_interval_op, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g.stream, op,
                     g.scratch_in.as<IV>(), b ? g.scratch_in2.as<IV>() : nullptr, g.scratch_out.as<IV>(),
                     g.verdicts.as<int>(), n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, g.scratch_out.p, bytes, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipMemcpyAsync(trap, g.verdicts.p, n * sizeof(int), hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  return CPL_OK;
}

int cpl_debug_prefilter(int family, const double* cells, size_t n, uint8_t* fp32_result, uint8_t* verdicts) {
  int rc = ensure_init();
  if (rc) return rc;
  if (family < 0 || family >= CPL_NUM_FAMILIES || !cells || !fp32_result || !verdicts)
    return fail(CPL_ERR_ARG, "cpl_debug_prefilter: bad argument");
  if (n == 0) return CPL_OK;
  if ((rc = g.cells.ensure(8 * n * sizeof(double))) || (rc = g.verdicts.ensure(2 * n))) return rc;
  HIP_TRY(hipMemcpyAsync(g.cells.p, cells, 8 * n * sizeof(double), hipMemcpyHostToDevice, g.stream));
  dispatch_family<LaunchPrefilterDebug>(family, (unsigned)((n + 255) / 256), g.cells.as<double>(), n,
                                        g.verdicts.as<uint8_t>(), g.verdicts.as<uint8_t>() + n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(fp32_result, g.verdicts.p, n, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipMemcpyAsync(verdicts, g.verdicts.as<uint8_t>() + n, n, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  return CPL_OK;
}

}  // extern "C"

namespace {

// Everything a refinement run derives from its configuration.
struct Plan {
  std::vector<int> fams;
  int nf = 0;
  int n_children = 2;
  int la_levels = 1;
  uint64_t n_roots = 1;
  uint64_t n_triples = 1;
  uint64_t la_entries = 0;
  uint64_t cap = 0;       // items per family region
  uint32_t seg_cap = 0;   // items per segment: cap / kSegs
  bool digest = true;
  bool flags = false;
  bool ext = false;       // extended kernels (FP32 prefilter)
  bool learned = false;
  bool last_round_proposals = false;  // the last round reads proposals (gate or check order; not the split axis)
  uint32_t prefilter_depths = 0;
  int block = 256;
  int cpt = 1;
  size_t stats_bytes = 0;
  uint64_t flag_words = 0;
};

int make_plan(const cpl_refine_cfg* cfg, bool want_flags, Plan& pl) {
  if (cfg->max_depth < 0 || cfg->max_depth >= CPL_MAX_ROUNDS) return fail(CPL_ERR_ARG, "max_depth out of range");
  pl.n_roots = 1;
  for (int a = 0; a < 4; ++a) {
    if (cfg->grid[a] < 1) return fail(CPL_ERR_ARG, "grid counts must be >= 1");
    pl.n_roots *= (uint64_t)cfg->grid[a];
  }
  if (pl.n_roots > 0xFFFFFFFFull) return fail(CPL_ERR_ARG, "more than 2^32 root cells");
  pl.n_triples = pl.n_roots / (uint64_t)cfg->grid[3];
  if (cfg->split_axes == 0 || (cfg->split_axes & ~0xFu)) return fail(CPL_ERR_ARG, "split_axes must be a nonzero subset of 0xF");
  for (int f = 0; f < CPL_NUM_FAMILIES; ++f)
    if (cfg->family_mask & (1u << f)) pl.fams.push_back(f);
  if (pl.fams.empty()) return fail(CPL_ERR_ARG, "empty family mask");
  pl.nf = (int)pl.fams.size();
  pl.n_children = cfg->split_mode == 0 ? 2 : (1 << __builtin_popcount(cfg->split_axes));
  pl.la_levels = 1 + ((cfg->split_axes & 1) ? cfg->max_depth : 0);
  pl.la_entries = (uint64_t)cfg->grid[0] * ((1ull << pl.la_levels) - 1);
  const char* cap_env = getenv("CPL_WORKLIST_CAP");
  pl.cap = cap_env ? strtoull(cap_env, nullptr, 10) : (1ull << 23);
  if (pl.cap < (uint64_t)kSegs || pl.cap > (1ull << 31))
    return fail(CPL_ERR_ARG, "CPL_WORKLIST_CAP must be between 64 and 2^31 items");
  pl.seg_cap = (uint32_t)(pl.cap / kSegs);
  pl.digest = !(cfg->options & CPL_OPT_NO_DIGEST);
  pl.flags = want_flags;
  pl.learned = (cfg->options & (CPL_OPT_LEARNED | CPL_OPT_LEARNED_SPLIT | CPL_OPT_LEARNED_ORDER)) != 0;
  pl.last_round_proposals = (cfg->options & (CPL_OPT_LEARNED | CPL_OPT_LEARNED_ORDER)) != 0;
  pl.ext = (cfg->options & CPL_OPT_FP32) != 0;
  pl.prefilter_depths = !(cfg->options & CPL_OPT_FP32) ? 0u : cfg->prefilter_depths ? cfg->prefilter_depths : 0xFFFFFFFFu;
  if (pl.learned)
    for (int f : pl.fams)
      if (!g.mlp_ready[f]) return fail(CPL_ERR_ARG, "learned options need a proposal installed for every family (cpl_mlp_install)");
  pl.block = cfg->block > 0 ? cfg->block : 256;
  if (pl.block % 64 || pl.block > 1024) return fail(CPL_ERR_ARG, "block must be a multiple of 64 up to 1024");
  pl.cpt = cfg->cells_per_thread > 0 ? cfg->cells_per_thread : 1;
  pl.stats_bytes = (size_t)CPL_MAX_ROUNDS * CPL_NUM_FAMILIES * STAT_SLOTS * sizeof(RoundStats);
  pl.flag_words = (pl.n_roots * pl.nf + 3) / 4;
  return CPL_OK;
}

// Counter buffer: the segment counts of five worklist regions per family slot
// (sets 0, 1, 4: the round regions of the graph path's rotation, set 2: live
// root items, set 3 unused), then one overflow word per family slot and the
// plain path's item totals.
constexpr size_t kSegSetWords = (size_t)kSegs * kSegStride;  // u64 words per region
constexpr size_t kSegCountBytes = 5 * CPL_NUM_FAMILIES * kSegSetWords * sizeof(unsigned long long);
constexpr size_t kCounterBytes = kSegCountBytes + 256;

// Device buffers: worklists [0], [1] (ping-pong) and [2] (live root items),
// each with one region per family; counters[5][family slots] + one overflow
// word per family slot; one la table per family slot (the graph form's chains
// each compute their own; the other forms use slot 0's).
int ensure_buffers(const Plan& pl) {
  int rc;
  const size_t per_item = 8 * sizeof(double) + 3 * sizeof(uint32_t);
  if ((rc = g.la_table.ensure(pl.la_entries * 3 * sizeof(double) * CPL_NUM_FAMILIES))) return rc;
  if ((rc = g.root_flags.ensure(pl.flag_words * 4 + 4))) return rc;
  if ((rc = g.stats.ensure(pl.stats_bytes))) return rc;
  if ((rc = g.stats_sum.ensure(CPL_MAX_ROUNDS * CPL_NUM_FAMILIES * sizeof(RoundStats)))) return rc;
  if ((rc = g.counters.ensure(kCounterBytes))) return rc;
  for (int b = 0; b < 2; ++b)
    if ((rc = g.wl[b].ensure((size_t)pl.cap * per_item * pl.nf))) return rc;
  if (pl.learned && (rc = g.prop.ensure((size_t)pl.cap * pl.nf))) return rc;
  g.wl_cap_items = pl.cap;
  return CPL_OK;
}

unsigned long long* counter(int buf, int slot) {
  return g.counters.as<unsigned long long>() + ((size_t)buf * CPL_NUM_FAMILIES + slot) * kSegSetWords;
}
uint32_t* overflow_word(int slot = 0) { return reinterpret_cast<uint32_t*>(g.counters.as<char>() + kSegCountBytes) + slot; }
uint64_t* item_totals() { return reinterpret_cast<uint64_t*>(g.counters.as<char>() + kSegCountBytes + 64); }
double* la_copy(const Plan& pl, int slot) { return g.la_table.as<double>() + (size_t)slot * pl.la_entries * 3; }
// Round counters of the graph path: round d reads rot_counter(d) and writes
// rot_counter(d + 1), and clears rot_counter(d + 2) for the next round (it was
// the input count of round d - 1), so no clearing node runs between rounds.
unsigned long long* rot_counter(int d, int slot) {
  static const int set[3] = {0, 1, 4};
  return counter(set[d % 3], slot);
}

RefineParams base_params(const cpl_refine_cfg* cfg, const Plan& pl) {
  RefineParams P;
  std::memset(&P, 0, sizeof(P));
  for (int a = 0; a < 4; ++a) {
    P.grid[a] = cfg->grid[a];
    P.box_lb[a] = cfg->box_lb[a];
    P.box_ub[a] = cfg->box_ub[a];
  }
  P.clip = cfg->clip;
  P.max_depth = cfg->max_depth;
  P.split_mode = cfg->split_mode;
  P.split_axes = cfg->split_axes;
  P.n_children = pl.n_children;
  P.nfam = pl.nf;
  P.n_roots = pl.n_roots;
  P.la_table = g.la_table.as<double>();
  P.root_flags = pl.flags ? g.root_flags.as<uint32_t>() : nullptr;
  P.stats = g.stats.as<RoundStats>();
  P.overflow = overflow_word();
  P.seg_cap = pl.seg_cap;
  P.digest = pl.digest ? 1 : 0;
  P.prefilter_depths = pl.prefilter_depths;
  P.mlp_threshold = cfg->learned_threshold;
  P.learned_split = (cfg->options & CPL_OPT_LEARNED_SPLIT) ? 1 : 0;
  P.learned_order = (cfg->options & CPL_OPT_LEARNED_ORDER) ? 1 : 0;
  return P;
}

// Per-launch family fields: slot and learned proposal.
void set_family(RefineParams& P, const Plan& pl, int s) {
  P.fam_slot = s;
  P.mlp = pl.learned ? g.mlp.as<MlpPacked>() + pl.fams[s] : nullptr;
}

// Proposal bytes of the items of `in` (n, or the counts P.in_seg when set) into the
// family's proposal region; the round launched next reads them through P.prop.
// n_expect: the round's item count (or its expected count for device-sized
// rounds); the grid gives each wave about kProposeGroups 64-cell groups.
constexpr uint64_t kProposeGroups = 8;
void launch_propose(const Plan& pl, int s, RefineParams& P, const Worklist& in, uint64_t n, uint64_t n_expect) {
  P.prop = nullptr;
  if (!pl.learned) return;
  // the last round splits nothing: with the split axis the only use of the
  // proposal, its round needs none
  if (P.depth >= P.max_depth && !pl.last_round_proposals) return;
  uint8_t* out = g.prop.as<uint8_t>() + (size_t)s * pl.cap;
  const uint64_t per_block = 256 * kProposeGroups;
  const uint64_t cap = (uint64_t)std::max(g.cu_count, 1) * 8;
  const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min((std::max<uint64_t>(n_expect, 1) + per_block - 1) / per_block, cap));
  hipLaunchKernelGGL(k_propose, dim3(grid), dim3(256), 0, launch_stream(), P, in, n, out);
  P.prop = out;
}

void fill(void* p, uint64_t words, uint32_t v) {
  unsigned grid = (unsigned)std::min<uint64_t>((words + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(std::max(grid, 1u)), dim3(256), 0, launch_stream(), static_cast<uint32_t*>(p), words, v);
}

template <int F> struct LaunchTriples {
  static void run(unsigned grid, int block, const RefineParams& P) {
    hipLaunchKernelGGL(k_root_triples<F>, dim3(grid), dim3(block), 0, launch_stream(), P);
  }
};

// Root items the triple screening passed on, per family, in the last collected
// graph-form run (sizes the next capture's depth-0 round).
uint64_t g_live_items[CPL_NUM_FAMILIES];

int ensure_stage() {
  if (!g.stage) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.stage), kStageBytes, hipHostMallocDefault));
  return CPL_OK;
}

// Sums the statistics copies on the device and queues their copy into the
// pinned stage (read after the stream has been synchronised).
int collect_enqueue() {
  const int words = CPL_MAX_ROUNDS * CPL_NUM_FAMILIES * (int)(sizeof(RoundStats) / 8);
  hipLaunchKernelGGL(k_reduce_stats, dim3((words + 255) / 256), dim3(256), 0, g.stream, g.stats.as<RoundStats>(),
                     g.stats_sum.as<RoundStats>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(g.stage, g.stats_sum.p, kStageStats, hipMemcpyDeviceToHost, g.stream));
  return CPL_OK;
}

int collect_staged(const Plan& pl, cpl_refine_stats* stats, uint8_t* root_flags);

int collect(const Plan& pl, cpl_refine_stats* stats, uint8_t* root_flags) {
  int rc = ensure_stage();
  if (rc) return rc;
  rc = collect_enqueue();
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g.stream));
  return collect_staged(pl, stats, root_flags);
}

// The rest of collect() once the stage holds the summed records.
int collect_staged(const Plan& pl, cpl_refine_stats* stats, uint8_t* root_flags) {
  std::vector<RoundStats> rs(CPL_MAX_ROUNDS * CPL_NUM_FAMILIES);
  std::memcpy(rs.data(), g.stage, kStageStats);
  if (root_flags) HIP_TRY(hipMemcpy(root_flags, g.root_flags.p, pl.n_roots * pl.nf, hipMemcpyDeviceToHost));
  double seconds = stats->seconds, eval_seconds = stats->eval_seconds;
  uint64_t peak = stats->peak_items;
  double launch[CPL_MAX_ROUNDS][CPL_NUM_FAMILIES];
  std::memcpy(launch, stats->launch_seconds, sizeof(launch));
  std::memset(stats, 0, sizeof(*stats));
  std::memcpy(stats->launch_seconds, launch, sizeof(launch));
  bool planned[CPL_NUM_FAMILIES] = {};
  for (int f : pl.fams) planned[f] = true;
  for (int d = 0; d < CPL_MAX_ROUNDS; ++d) {
    for (int f = 0; f < CPL_NUM_FAMILIES; ++f) {
      if (!planned[f]) continue;  // records of other families are not cleared by every form
      const RoundStats& r = rs[d * CPL_NUM_FAMILIES + f];
      stats->evaluations[d] += r.evaluations;
      for (int v = 0; v < 8; ++v) {
        stats->verdicts[d][v] += r.verdicts[v];
        stats->family_verdicts[d][f][v] = r.verdicts[v];
      }
      stats->digest[d] += r.digest;
      stats->verdict_digest[d] += r.vdigest;
      stats->certified += r.certified;
      stats->unresolved += r.unresolved;
      stats->fp32_settled[d][f] = r.fp32_settled;
      if (d == 0) g_live_items[f] = r.live_items;
      stats->fp32_tried[d][f] = r.fp32_tried;
      stats->total_evaluations += r.evaluations;
    }
    if (stats->evaluations[d]) stats->rounds = d + 1;
  }
  stats->seconds = seconds;
  stats->eval_seconds = eval_seconds;
  stats->peak_items = peak;
  return CPL_OK;
}

// The plain path: one launch per family and round, the host reads the next
// round's sizes after each round (sizes the grids, detects overflow), HIP
// events time every launch.
int refine_plain(const cpl_refine_cfg* cfg, const Plan& pl, cpl_refine_stats* stats, uint8_t* root_flags) {
  hipEvent_t ev0, ev1;
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  std::vector<hipEvent_t> kev;
  std::vector<int> kslot;
  if (pl.flags) HIP_TRY(hipMemsetAsync(g.root_flags.p, 0, pl.flag_words * 4, g.stream));
  HIP_TRY(hipMemsetAsync(g.stats.p, 0, pl.stats_bytes, g.stream));
  HIP_TRY(hipMemsetAsync(g.counters.p, 0, kCounterBytes, g.stream));
  HIP_TRY(hipEventRecord(ev0, g.stream));
  hipLaunchKernelGGL(k_la_table, dim3((unsigned)((pl.la_entries + 255) / 256)), dim3(256), 0, g.stream, cfg->box_lb[0],
                     cfg->box_ub[0], cfg->grid[0], pl.la_levels, cfg->clip, g.la_table.as<double>());
  HIP_TRY(hipGetLastError());
  RefineParams P = base_params(cfg, pl);
  std::vector<uint64_t> cur_n(pl.nf, 0);
  uint64_t peak = 0;
  int cur = 0;
  for (int d = 0; d <= cfg->max_depth; ++d) {
    const int nxt = cur ^ 1;
    P.depth = d;
    bool any = false;
    for (int s = 0; s < pl.nf; ++s) {
      uint64_t n = d == 0 ? pl.n_roots : cur_n[s];
      if (n == 0) continue;
      any = true;
      set_family(P, pl, s);
      P.out = region(g.wl[nxt], pl.cap, s);
      P.out_seg = counter(nxt, s);
      P.in_seg = d == 0 ? nullptr : counter(cur, s);
      Worklist in = region(g.wl[cur], pl.cap, s);
      hipEvent_t a, b;
      HIP_TRY(hipEventCreate(&a));
      HIP_TRY(hipEventCreate(&b));
      if (g.recording) {
        dispatch_family<LaunchRecord>(pl.fams[s], d == 0, grid_for(n, 256), P, in, n, g.sink);
        HIP_TRY(hipGetLastError());
      }
      P.prop = nullptr;
      if (d > 0) launch_propose(pl, s, P, in, n, n);
      HIP_TRY(hipEventRecord(a, g.stream));
      dispatch_family<LaunchRound>(pl.fams[s], d == 0, pl.ext, grid_for(n, 256), P, in, n);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipEventRecord(b, g.stream));
      kev.push_back(a);
      kev.push_back(b);
      kslot.push_back(d * CPL_NUM_FAMILIES + pl.fams[s]);
    }
    if (!any) break;
    uint64_t h[CPL_NUM_FAMILIES];
    uint32_t ovf = 0;
    hipLaunchKernelGGL(k_seg_totals, dim3(pl.nf), dim3(64), 0, g.stream, counter(nxt, 0), pl.seg_cap, item_totals());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h, item_totals(), pl.nf * sizeof(uint64_t), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(&ovf, overflow_word(), sizeof(uint32_t), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemsetAsync(counter(cur, 0), 0, CPL_NUM_FAMILIES * kSegSetWords * sizeof(unsigned long long), g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (ovf) return fail(CPL_ERR_CAPACITY, "worklist capacity exceeded; raise CPL_WORKLIST_CAP");
    uint64_t tot = 0;
    for (int s = 0; s < pl.nf; ++s) {
      cur_n[s] = h[s];
      tot += h[s];
    }
    peak = std::max(peak, tot);
    cur = nxt;
  }
  HIP_TRY(hipEventRecord(ev1, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  std::memset(stats, 0, sizeof(*stats));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  stats->seconds = ms * 1e-3;
  double eval_ms = 0.0;
  for (size_t k = 0; k + 1 < kev.size(); k += 2) {
    float m = 0.f;
    HIP_TRY(hipEventElapsedTime(&m, kev[k], kev[k + 1]));
    eval_ms += m;
    stats->launch_seconds[kslot[k / 2] / CPL_NUM_FAMILIES][kslot[k / 2] % CPL_NUM_FAMILIES] = m * 1e-3;
    (void)hipEventDestroy(kev[k]);
    (void)hipEventDestroy(kev[k + 1]);
  }
  stats->eval_seconds = eval_ms * 1e-3;
  stats->peak_items = peak;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  return collect(pl, stats, root_flags);
}

// Fork: every family branch waits for the work enqueued so far on the library
// stream; join: the library stream waits for every branch.  Inside a stream
// capture this yields a graph with one independent chain per family.
// Stream of family slot s: slot 0 runs on the library stream itself, so four
// families use four streams (the hardware offers four queues per process by
// default; a fifth stream would share a queue and serialise two families).
hipStream_t fam_stream_of(int s) { return s == 0 ? g.stream : g.fam_stream[s - 1]; }

void fork_families(int nf) {
  (void)hipEventRecord(g.fork_ev, g.stream);
  for (int s = 1; s < nf; ++s) (void)hipStreamWaitEvent(fam_stream_of(s), g.fork_ev, 0);
}
void join_families(int nf) {
  for (int s = 1; s < nf; ++s) {
    (void)hipEventRecord(g.join_ev[s], fam_stream_of(s));
    (void)hipStreamWaitEvent(g.stream, g.join_ev[s], 0);
  }
  g_launch_stream = nullptr;
}

// ---- graph path ---------------------------------------------------------------
// Captured graphs: the prologue (library stream) and one linear chain per
// family slot (launched on that family's stream, so the chains run
// concurrently; a single graph with parallel branches started them one after
// another).
struct GraphCache {
  bool valid = false;
  cpl_refine_cfg cfg;
  bool flags = false;
  hipGraphExec_t exec[1 + CPL_NUM_FAMILIES] = {};
  hipGraph_t graph[1 + CPL_NUM_FAMILIES] = {};
  int n = 0;
  // expected items per (round, slot) from the previous replay: sizes the grids;
  // then, per slot, the live root items of the triple screening
  std::vector<uint32_t> expect;
  // end of each family chain in the previous replay (ms after its start; 0: not measured)
  float chain_ms[CPL_NUM_FAMILIES] = {};
  // device buffers the captured kernels address: a buffer that another run
  // reallocated (or cpl_shutdown released) forces a new capture
  std::vector<void*> bufs;
  uint64_t cap = 0;  // worklist capacity per family region (CPL_WORKLIST_CAP) the grids and clamps were built for
  void reset() {
    for (int k = 0; k < 1 + CPL_NUM_FAMILIES; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graph[k]) (void)hipGraphDestroy(graph[k]);
      exec[k] = nullptr;
      graph[k] = nullptr;
    }
    n = 0;
    valid = false;
  }
};
GraphCache g_graph;

std::vector<void*> graph_buffers() {
  return {g.wl[0].p, g.wl[1].p, g.la_table.p, g.root_flags.p, g.stats.p, g.stats_sum.p, g.counters.p, g.prop.p, g.mlp.p};
}

void drop_graphs() {
  g_graph.reset();
  g_graph.bufs.clear();
}


unsigned grid_tuned(uint64_t expect, const Plan& pl) {
  uint64_t per_block = (uint64_t)pl.block * pl.cpt;
  uint64_t want = (std::max<uint64_t>(expect, 1) + per_block - 1) / per_block;
  uint64_t cap = (uint64_t)std::max(g.cu_count, 1) * 64;
  return (unsigned)std::max<uint64_t>(1, std::min(want, cap));
}

// Work shared by the families: clearing the per-root flags (their bytes
// interleave the families), enqueued on the library stream.  Only captured and
// replayed when flags are requested; then the family chains wait for it.
void enqueue_prologue(const Plan& pl) {
  g_launch_stream = nullptr;
  fill(g.root_flags.p, pl.flag_words, 0u);
}

// Round d of family slot s with device-side sizes, enqueued on that family's
// stream (a family's rounds form one linear chain; families are independent).
void enqueue_round(const cpl_refine_cfg* cfg, const Plan& pl, int s, int d, const std::vector<uint32_t>& expect) {
  const bool hier = (cfg->options & CPL_OPT_HIER_ROOTS) != 0;
  RefineParams P = base_params(cfg, pl);
  auto exp_of = [&](int dd) -> uint64_t {
    size_t k = (size_t)dd * CPL_NUM_FAMILIES + s;
    return k < expect.size() && expect[k] ? expect[k] : pl.cap;
  };
  // items of the depth-0 round after the triple screening (all of round 0's
  // evaluations before the first replay)
  auto live_of = [&]() -> uint64_t {
    size_t k = (size_t)CPL_MAX_ROUNDS * CPL_NUM_FAMILIES + s;
    return k < expect.size() && expect[k] ? expect[k] : exp_of(0);
  };
  g_launch_stream = fam_stream_of(s);
  set_family(P, pl, s);
  P.la_table = la_copy(pl, s);
  P.overflow = overflow_word(s);
  P.depth = d;
  if (d == 0) {
    P.in_seg = nullptr;
    P.out = region(g.wl[1], pl.cap, s);
    P.out_seg = rot_counter(1, s);
    P.zero_seg = rot_counter(2, s);
    P.prop = nullptr;  // the root round runs without proposals in every form
    if (hier) {
      // live root items go to buffer 0, evaluated from there into buffer 1
      RefineParams T = P;
      T.out = region(g.wl[0], pl.cap, s);
      T.out_seg = counter(2, s);
      T.zero_seg = nullptr;
      dispatch_family<LaunchTriples>(pl.fams[s], grid_tuned(pl.n_triples, pl), pl.block, T);
      P.in_seg = counter(2, s);
      dispatch_family<LaunchRoundB>(pl.fams[s], false, pl.ext, grid_tuned(live_of(), pl), pl.block, P,
                                    region(g.wl[0], pl.cap, s), (uint64_t)0);
    } else {
      Worklist none{};
      dispatch_family<LaunchRoundB>(pl.fams[s], true, pl.ext, grid_tuned(pl.n_roots, pl), pl.block, P, none,
                                    pl.n_roots);
    }
  } else {
    // round d reads buffer 1 - d % 2 ... rounds alternate buffers 1 -> 0 -> 1
    const int cur = d % 2 == 1 ? 1 : 0, nxt = cur ^ 1;
    P.in_seg = rot_counter(d, s);
    P.out = region(g.wl[nxt], pl.cap, s);
    P.out_seg = rot_counter(d + 1, s);
    P.zero_seg = rot_counter(d + 2, s);
    launch_propose(pl, s, P, region(g.wl[cur], pl.cap, s), 0, exp_of(d));
    dispatch_family<LaunchRoundB>(pl.fams[s], false, pl.ext, grid_tuned(exp_of(d), pl), pl.block, P,
                                  region(g.wl[cur], pl.cap, s), (uint64_t)0);
  }
  g_launch_stream = nullptr;
}

void enqueue_family(const cpl_refine_cfg* cfg, const Plan& pl, int s, const std::vector<uint32_t>& expect) {
  g_launch_stream = fam_stream_of(s);
  const int words = CPL_MAX_ROUNDS * STAT_SLOTS * (int)(sizeof(RoundStats) / sizeof(unsigned long long));
  hipLaunchKernelGGL(k_clear_family, dim3((words + 255) / 256), dim3(256), 0, launch_stream(), g.stats.as<RoundStats>(),
                     pl.fams[s], counter(0, 0), overflow_word(0), s);
  hipLaunchKernelGGL(k_la_table, dim3((unsigned)((pl.la_entries + 255) / 256)), dim3(256), 0, launch_stream(),
                     cfg->box_lb[0], cfg->box_ub[0], cfg->grid[0], pl.la_levels, cfg->clip, la_copy(pl, s));
  g_launch_stream = nullptr;
  for (int d = 0; d <= cfg->max_depth; ++d) enqueue_round(cfg, pl, s, d, expect);
}

int refine_graph(const cpl_refine_cfg* cfg, const Plan& pl, cpl_refine_stats* stats, uint8_t* root_flags) {
  bool same = g_graph.valid && g_graph.flags == pl.flags && std::memcmp(&g_graph.cfg, cfg, sizeof(*cfg)) == 0 &&
              g_graph.cap == pl.cap && g_graph.bufs == graph_buffers();
  if (!same) {
    // chain times of another configuration do not order this one's chains
    if (std::memcmp(&g_graph.cfg, cfg, sizeof(*cfg)) != 0)
      for (float& m : g_graph.chain_ms) m = 0.f;
    g_graph.reset();
    for (int k = 0; k <= pl.nf; ++k) {
      hipStream_t cs = k == 0 ? g.stream : fam_stream_of(k - 1);
      if (k == 0 && !pl.flags) continue;  // no shared prologue
      HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      if (k == 0) enqueue_prologue(pl);
      else enqueue_family(cfg, pl, k - 1, g_graph.expect);
      hipError_t e = hipStreamEndCapture(cs, &g_graph.graph[k]);
      if (e != hipSuccess) return fail(CPL_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      HIP_TRY(hipGraphInstantiate(&g_graph.exec[k], g_graph.graph[k], nullptr, nullptr, 0));
    }
    g_graph.n = pl.nf;
    g_graph.cfg = *cfg;
    g_graph.flags = pl.flags;
    g_graph.bufs = graph_buffers();
    g_graph.cap = pl.cap;
    g_graph.valid = true;
  }
  int rc = ensure_stage();
  if (rc) return rc;
  hipEvent_t ev0 = g.run_begin, ev1 = g.run_end;
  HIP_TRY(hipEventRecord(ev0, g.stream));
  const bool prologue = g_graph.exec[0] != nullptr;
  if (prologue) {
    HIP_TRY(hipGraphLaunch(g_graph.exec[0], g.stream));
    HIP_TRY(hipEventRecord(g.fork_ev, g.stream));
  }
  // Submit first the chain that ended last in the previous replay (host
  // submission of a captured chain takes tens of microseconds, so later chains
  // start later); before any replay, the chain with the most work: the items of
  // its refinement rounds and its live root items (round 0's out-of-domain
  // roots, screened per triple, cost little).
  std::vector<int> order(pl.nf);
  std::vector<double> work(pl.nf, 0.0);
  bool timed = true;
  for (int s = 0; s < pl.nf; ++s) timed = timed && g_graph.chain_ms[s] > 0.f;
  for (int s = 0; s < pl.nf; ++s) {
    order[s] = s;
    if (timed) {
      work[s] = g_graph.chain_ms[s];
      continue;
    }
    for (int d = 1; d <= CPL_MAX_ROUNDS; ++d) {
      const size_t k = (size_t)d * CPL_NUM_FAMILIES + s;
      if (k < g_graph.expect.size()) work[s] += g_graph.expect[k];
    }
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  for (int s : order) {
    if (s == 0) {
      HIP_TRY(hipGraphLaunch(g_graph.exec[1], g.stream));
      HIP_TRY(hipEventRecord(g.chain_end[0], g.stream));
      continue;
    }
    if (prologue) HIP_TRY(hipStreamWaitEvent(fam_stream_of(s), g.fork_ev, 0));
    HIP_TRY(hipGraphLaunch(g_graph.exec[1 + s], fam_stream_of(s)));
    HIP_TRY(hipEventRecord(g.chain_end[s], fam_stream_of(s)));
    HIP_TRY(hipEventRecord(g.join_ev[s], fam_stream_of(s)));
  }
  for (int s = 1; s < pl.nf; ++s) HIP_TRY(hipStreamWaitEvent(g.stream, g.join_ev[s], 0));
  HIP_TRY(hipEventRecord(ev1, g.stream));
  // statistics and overflow words come back with the replay's one synchronisation
  rc = collect_enqueue();
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(g.stage + kStageStats, overflow_word(0), pl.nf * sizeof(uint32_t), hipMemcpyDeviceToHost,
                         g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  for (int s = 0; s < pl.nf; ++s) {
    float m = 0.f;
    HIP_TRY(hipEventElapsedTime(&m, ev0, g.chain_end[s]));
    g_graph.chain_ms[s] = m;
  }
  uint32_t ovfw[CPL_NUM_FAMILIES] = {};
  std::memcpy(ovfw, g.stage + kStageStats, pl.nf * sizeof(uint32_t));
  uint32_t ovf = 0;
  for (int s = 0; s < pl.nf; ++s) ovf |= ovfw[s];
  std::memset(stats, 0, sizeof(*stats));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  stats->seconds = ms * 1e-3;
  stats->eval_seconds = stats->seconds;
  if (ovf) {
    g_graph.reset();
    return fail(CPL_ERR_CAPACITY, "worklist capacity exceeded; raise CPL_WORKLIST_CAP");
  }
  rc = collect_staged(pl, stats, root_flags);
  if (rc) return rc;
  // Record the round sizes of this replay; if they differ from the ones the
  // graph's grids were sized for, re-capture on the next call.
  std::vector<uint32_t> seen((size_t)(CPL_MAX_ROUNDS + 1) * CPL_NUM_FAMILIES, 0);
  for (int d = 0; d < stats->rounds; ++d)
    for (int s = 0; s < pl.nf; ++s) {
      uint64_t n = 0;
      for (int v = 0; v < 8; ++v) n += stats->family_verdicts[d][pl.fams[s]][v];
      seen[(size_t)d * CPL_NUM_FAMILIES + s] = (uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFu);
    }
  if (cfg->options & CPL_OPT_HIER_ROOTS)
    for (int s = 0; s < pl.nf; ++s)
      seen[(size_t)CPL_MAX_ROUNDS * CPL_NUM_FAMILIES + s] =
          (uint32_t)std::min<uint64_t>(g_live_items[pl.fams[s]], 0xFFFFFFFFu);
  if (seen != g_graph.expect) {
    g_graph.expect = seen;
    g_graph.valid = false;  // keep the exec until the next capture replaces it
  }
  return CPL_OK;
}

// ---- persistent path ---------------------------------------------------------
template <int F> struct LaunchPersist {
  static const void* fn(bool p32) {
    return p32 ? reinterpret_cast<const void*>(k_persist<F, true>) : reinterpret_cast<const void*>(k_persist<F, false>);
  }
  static int grid_for_lds(size_t lds, bool p32) {
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn(p32), 64, lds);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb * std::max(g.cu_count, 1);
  }
  static void run(int grid, size_t lds, bool p32, const PersistParams& Q) {
    if (p32) hipLaunchKernelGGL((k_persist<F, true>), dim3(grid), dim3(64), lds, launch_stream(), Q);
    else hipLaunchKernelGGL((k_persist<F, false>), dim3(grid), dim3(64), lds, launch_stream(), Q);
  }
};

int refine_persistent(const cpl_refine_cfg* cfg, const Plan& pl, cpl_refine_stats* stats, uint8_t* root_flags) {
  int rc;
  const int nch = pl.n_children;
  // LDS stack of each 64-lane block: after a refill the stack holds at most 127
  // cells (fewer than 64, plus 64 from the spill stack), a batch pops 64 and
  // pushes at most 64 * nch, and cells beyond CAP - 64 * nch are spilled, so
  // CAP = 64 + 64 * nch is the least that never overflows; the smaller stack
  // leaves room for more blocks per CU (2.70-2.81 against 2.76-3.09 ms per
  // proof with 128 + 64 * nch)
  const int cap_lds = 64 + 64 * nch;
  const size_t lds = (size_t)cap_lds * (8 * sizeof(double) + 3 * sizeof(uint32_t)) +
                     (size_t)CPL_MAX_ROUNDS * ST_N * sizeof(unsigned long long);
  const size_t per_item = 8 * sizeof(double) + 3 * sizeof(uint32_t);
  int grid_max = 0;
  for (int f : pl.fams) {
    int gsz = 0;
    switch (f) {
      case 0: gsz = LaunchPersist<0>::grid_for_lds(lds, pl.ext); break;
      case 1: gsz = LaunchPersist<1>::grid_for_lds(lds, pl.ext); break;
      case 2: gsz = LaunchPersist<2>::grid_for_lds(lds, pl.ext); break;
      case 3: gsz = LaunchPersist<3>::grid_for_lds(lds, pl.ext); break;
      case 4: gsz = LaunchPersist<4>::grid_for_lds(lds, pl.ext); break;
      default: gsz = LaunchPersist<5>::grid_for_lds(lds, pl.ext); break;
    }
    grid_max = std::max(grid_max, gsz);
  }
  // private spill stacks live in worklist buffer 1 (unused by this path), one
  // partition per family because the families' kernels run concurrently
  uint32_t spill_cap =
      (uint32_t)std::min<uint64_t>(g.wl[1].bytes / ((size_t)pl.nf * grid_max * per_item) / 64 * 64, 1u << 20);
  if (spill_cap < 256) return fail(CPL_ERR_CAPACITY, "persistent form: worklist buffer too small for the spill stacks");
  if ((rc = g.persist_q.ensure((size_t)CPL_NUM_FAMILIES * 8 * sizeof(uint32_t)))) return rc;

  hipEvent_t ev0, ev1;
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  if (pl.flags) HIP_TRY(hipMemsetAsync(g.root_flags.p, 0, pl.flag_words * 4, g.stream));
  HIP_TRY(hipMemsetAsync(g.stats.p, 0, pl.stats_bytes, g.stream));
  HIP_TRY(hipMemsetAsync(g.counters.p, 0, kCounterBytes, g.stream));
  HIP_TRY(hipMemsetAsync(g.persist_q.p, 0, (size_t)CPL_NUM_FAMILIES * 8 * sizeof(uint32_t), g.stream));
  HIP_TRY(hipEventRecord(ev0, g.stream));
  hipLaunchKernelGGL(k_la_table, dim3((unsigned)((pl.la_entries + 255) / 256)), dim3(256), 0, g.stream, cfg->box_lb[0],
                     cfg->box_ub[0], cfg->grid[0], pl.la_levels, cfg->clip, g.la_table.as<double>());
  RefineParams P = base_params(cfg, pl);
  P.depth = 0;
  fork_families(pl.nf);
  for (int s = 0; s < pl.nf; ++s) {
    g_launch_stream = fam_stream_of(s);
    set_family(P, pl, s);
    RefineParams T = P;
    T.out = region(g.wl[0], pl.cap, s);
    T.out_seg = counter(2, s);
    dispatch_family<LaunchTriples>(pl.fams[s], grid_tuned(pl.n_triples, pl), pl.block, T);
    HIP_TRY(hipGetLastError());
    PersistParams Q;
    Q.R = P;
    Q.roots = region(g.wl[0], pl.cap, s);
    Q.root_seg = counter(2, s);
    Q.q = g.persist_q.as<uint32_t>() + 8 * s;
    char* part = g.wl[1].as<char>() + (size_t)s * grid_max * spill_cap * per_item;
    Q.sp_e = reinterpret_cast<double*>(part);
    Q.sp_u = reinterpret_cast<uint32_t*>(part + (size_t)grid_max * 8 * spill_cap * sizeof(double));
    Q.spill_cap = spill_cap;
    Q.cap = cap_lds;
    Q.claim = (uint32_t)std::min(64, 8 * pl.cpt);
    switch (pl.fams[s]) {
      case 0: LaunchPersist<0>::run(LaunchPersist<0>::grid_for_lds(lds, pl.ext), lds, pl.ext, Q); break;
      case 1: LaunchPersist<1>::run(LaunchPersist<1>::grid_for_lds(lds, pl.ext), lds, pl.ext, Q); break;
      case 2: LaunchPersist<2>::run(LaunchPersist<2>::grid_for_lds(lds, pl.ext), lds, pl.ext, Q); break;
      case 3: LaunchPersist<3>::run(LaunchPersist<3>::grid_for_lds(lds, pl.ext), lds, pl.ext, Q); break;
      case 4: LaunchPersist<4>::run(LaunchPersist<4>::grid_for_lds(lds, pl.ext), lds, pl.ext, Q); break;
      default: LaunchPersist<5>::run(LaunchPersist<5>::grid_for_lds(lds, pl.ext), lds, pl.ext, Q); break;
    }
    HIP_TRY(hipGetLastError());
  }
  join_families(pl.nf);
  HIP_TRY(hipEventRecord(ev1, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  std::vector<uint32_t> qs((size_t)CPL_NUM_FAMILIES * 8);
  HIP_TRY(hipMemcpy(qs.data(), g.persist_q.p, qs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  uint32_t ovf = 0;
  HIP_TRY(hipMemcpy(&ovf, overflow_word(), sizeof(ovf), hipMemcpyDeviceToHost));
  std::memset(stats, 0, sizeof(*stats));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  stats->seconds = ms * 1e-3;
  stats->eval_seconds = stats->seconds;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  for (int s = 0; s < pl.nf; ++s)
    if (qs[8 * s + Q_ERR]) return fail(CPL_ERR_CAPACITY, "persistent form: a private spill stack overflowed");
  if (ovf) return fail(CPL_ERR_CAPACITY, "root worklist capacity exceeded; raise CPL_WORKLIST_CAP");
  return collect(pl, stats, root_flags);
}

}  // namespace

extern "C" {

int cpl_refine_run(const cpl_refine_cfg* cfg, cpl_refine_stats* stats, uint8_t* root_flags) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!cfg || !stats) return fail(CPL_ERR_ARG, "cpl_refine_run: null cfg or stats");
  Plan pl;
  if ((rc = make_plan(cfg, root_flags != nullptr, pl))) return rc;
  if ((rc = ensure_buffers(pl))) return rc;
  const uint32_t known = CPL_OPT_GRAPH | CPL_OPT_PERSISTENT | CPL_OPT_HIER_ROOTS | CPL_OPT_FP32 | CPL_OPT_LEARNED |
                         CPL_OPT_LEARNED_SPLIT | CPL_OPT_LEARNED_ORDER | CPL_OPT_NO_DIGEST;
  if (cfg->options & ~known) return fail(CPL_ERR_ARG, "unsupported option bits");
  if ((cfg->options & CPL_OPT_PERSISTENT) && pl.learned)
    return fail(CPL_ERR_ARG, "the learned proposal runs with the round-based forms (plain or CPL_OPT_GRAPH)");
  if (cfg->options & CPL_OPT_PERSISTENT) return refine_persistent(cfg, pl, stats, root_flags);
  if (cfg->options & CPL_OPT_GRAPH) return refine_graph(cfg, pl, stats, root_flags);
  if (cfg->options & CPL_OPT_HIER_ROOTS) return fail(CPL_ERR_ARG, "CPL_OPT_HIER_ROOTS needs CPL_OPT_GRAPH");
  return refine_plain(cfg, pl, stats, root_flags);
}

int cpl_mlp_install(int family, const float* params) {
  int rc = ensure_init();
  if (rc) return rc;
  if (family < 0 || family >= CPL_NUM_FAMILIES) return fail(CPL_ERR_ARG, "cpl_mlp_install: bad family");
  if (!params) {
    g.mlp_ready[family] = false;
    return CPL_OK;
  }
  if ((rc = g.mlp.ensure(CPL_NUM_FAMILIES * sizeof(MlpPacked)))) return rc;
  std::vector<char> host(sizeof(MlpPacked));
  cpl_mlp_pack(params, host.data());
  HIP_TRY(hipMemcpyAsync(g.mlp.as<MlpPacked>() + family, host.data(), host.size(), hipMemcpyHostToDevice, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  g.mlp_ready[family] = true;
  return CPL_OK;
}

int cpl_mlp_collect(const cpl_refine_cfg* cfg, uint64_t max_records, uint32_t keep_every, float* features,
                    float* targets, uint64_t* n_records) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!cfg || !features || !targets || !n_records || max_records == 0 || max_records > (1ull << 28))
    return fail(CPL_ERR_ARG, "cpl_mlp_collect: bad argument");
  cpl_refine_cfg c = *cfg;
  c.options = CPL_OPT_NO_DIGEST;
  Plan pl;
  if ((rc = make_plan(&c, false, pl))) return rc;
  if ((rc = ensure_buffers(pl))) return rc;
  if ((rc = g.records.ensure(max_records * 16 * sizeof(float) + 256))) return rc;
  g.sink.feat = g.records.as<float>();
  g.sink.targ = g.records.as<float>() + max_records * 8;
  g.sink.count = reinterpret_cast<uint32_t*>(g.records.as<float>() + max_records * 16);
  g.sink.cap = max_records;
  g.sink.keep = keep_every > 0 ? keep_every : 1;
  HIP_TRY(hipMemsetAsync(g.sink.count, 0, sizeof(uint32_t), g.stream));
  g.recording = true;
  cpl_refine_stats st;
  rc = refine_plain(&c, pl, &st, nullptr);
  g.recording = false;
  if (rc) return rc;
  uint32_t cnt = 0;
  HIP_TRY(hipMemcpy(&cnt, g.sink.count, sizeof(uint32_t), hipMemcpyDeviceToHost));
  const uint64_t n = std::min<uint64_t>(cnt, max_records);
  HIP_TRY(hipMemcpy(features, g.sink.feat, n * 8 * sizeof(float), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(targets, g.sink.targ, n * 8 * sizeof(float), hipMemcpyDeviceToHost));
  *n_records = n;
  return CPL_OK;
}

int cpl_autotune(cpl_refine_cfg* cfg, int reps, cpl_tune_point* points, int* n_points) {
  if (!cfg || reps < 1) return fail(CPL_ERR_ARG, "cpl_autotune: bad argument");
  static const int blocks[] = {64, 128, 256, 512, 1024};
  static const int cpts[] = {1, 2, 4, 8};
  cpl_refine_stats st;
  double best = 1e30;
  int bb = 256, bc = 1, k = 0;
  for (int b : blocks) {
    for (int c : cpts) {
      cpl_refine_cfg t = *cfg;
      t.block = b;
      t.cells_per_thread = c;
      int rc = cpl_refine_run(&t, &st, nullptr);  // warm-up (captures the graph)
      if (rc) return rc;
      rc = cpl_refine_run(&t, &st, nullptr);      // re-capture with measured round sizes
      if (rc) return rc;
      std::vector<double> times;
      for (int r = 0; r < reps; ++r) {
        if ((rc = cpl_refine_run(&t, &st, nullptr))) return rc;
        times.push_back(st.seconds);
      }
      std::sort(times.begin(), times.end());
      double med = times[times.size() / 2];
      if (points && k < 20) points[k] = cpl_tune_point{b, c, med};
      ++k;
      if (med < best) { best = med; bb = b; bc = c; }
    }
  }
  cfg->block = bb;
  cfg->cells_per_thread = bc;
  if (n_points) *n_points = std::min(k, 20);
  return CPL_OK;
}

// Reference subdivision counts per family (la, r1, r2, r3).
static const int32_t kRefSubdiv[CPL_NUM_FAMILIES][4] = {
    {256, 256, 256, 64},     // start_configuration.cu:41-44
    {4096, 4096, 4096, 1},   // start_configuration.cu:51-53
    {256, 256, 256, 64},     // start_configuration.cu:41-44 (reused by the 3d kernel)
    {512, 256, 256, 64},     // middle_configuration.cu:39-42
    {4096, 1024, 1024, 1},   // end_configuration.cu:39-41
    {1024, 256, 128, 128},   // end_configuration.cu:46-49
};

static const char* kCritText[CPL_NUM_FAMILIES][3] = {
    {"", "Critical interval: Start configuration, case 1: ", "Critical interval: Start configuration, case 2: "},
    {"", "Critical interval: Start & end configuration, 2 disks: ", ""},
    {"", "Start & end configuration: Critical interval - ", ""},
    {"", "Critical interval: Middle configuration, case 1.1: ", "Critical interval: Middle configuration, case 2.1: "},
    {"", "End configuration (2 disks): Critical interval with r1 on the inside - ",
     "End configuration (2 disks): Critical interval with r1 on the outside - "},
    {"", "End configuration (3 disks): Critical interval with r1 on the inside - ",
     "End configuration (3 disks): Critical interval with r1 on the outside - "},
};

namespace {

// Subdivision counts of a launch (reference counts when subdiv is NULL).
int mode_r_subdiv(int family, const int32_t* subdiv, int32_t n[4]) {
  if (family < 0 || family >= CPL_NUM_FAMILIES) return fail(CPL_ERR_ARG, "cpl_first_problematic: bad family");
  for (int k = 0; k < 4; ++k) n[k] = subdiv ? subdiv[k] : kRefSubdiv[family][k];
  if (!(family == 0 || family == 2 || family == 3 || family == 5)) n[3] = 1;
  for (int k = 0; k < 4; ++k)
    if (n[k] < 1) return fail(CPL_ERR_ARG, "subdivision counts must be >= 1");
  return CPL_OK;
}

// Enqueues one reference-grid launch on the library stream: clear the scratch,
// per-lambda r1 ranges, the family's kernel.  ranges: n[0] x 4 doubles; misc:
// MODE_R_MISC_BYTES; crit: crit_cap records of 9 doubles (output mode).
int mode_r_enqueue(int family, cpl_iv outer_la, const int32_t n[4], bool out, bool exhaustive, double* ranges,
                   unsigned long long* misc, double* crit, uint32_t crit_cap) {
  ModeRParams P;
  P.outer = IV(outer_la.lb, outer_la.ub);
  for (int k = 0; k < 4; ++k) P.n[k] = n[k];
  P.r1_ranges = ranges;
  P.misc = misc;
  P.exhaustive = exhaustive ? 1 : 0;
  P.crit_out = crit;
  P.crit_cap = crit ? crit_cap : 0u;
  hipLaunchKernelGGL(k_mode_r_init, dim3(1), dim3(256), 0, g.stream, misc);
  hipLaunchKernelGGL(k_mode_r_ranges, dim3((unsigned)((n[0] + 255) / 256)), dim3(256), 0, g.stream, P.outer, n[0],
                     ranges);
  // 16 blocks of four waves per CU; waves take (la, r1) rows in turn
  const uint64_t rows = (uint64_t)n[0] * n[1];
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((rows + 3) / 4, (uint64_t)std::max(g.cu_count, 1) * 16));
  dispatch_family<LaunchModeR>(family, out, (unsigned)blocks, P);
  HIP_TRY(hipGetLastError());
  return CPL_OK;
}

// Reads a launch's scratch back (after the stream has been synchronised) and
// applies the failure rule: a trap or fit-check failure counts where its la.lb
// is below the first problematic la.lb (every cell there is reached by the
// reference), or anywhere in output mode.
int mode_r_finish(bool out, const unsigned long long* misc, double* first_out, cpl_mode_r_info* info) {
  const uint64_t v = misc[MR_FIRST];
  {
    uint64_t mask = ((v >> 63) - 1) | (1ull << 63);
    uint64_t b = v ^ mask;
    std::memcpy(first_out, &b, 8);
  }
  cpl_mode_r_info I;
  std::memset(&I, 0, sizeof I);
  for (int k = 0; k < MR_NSLOTS; ++k) {
    const unsigned long long* slot = misc + MR_SLOTS + MR_SLOT_STRIDE * k;
    I.cells_evaluated += slot[0];
    I.rows_skipped += slot[1];
    I.nofit_flags += slot[2];
  }
  I.lines_found = out ? misc[MR_NREC] : 0;
  const uint64_t limit = out ? ~0ull : v;
  int code = 0;
  for (int c = 1; c < 6; ++c)
    if (misc[MR_FAIL + c] < limit && (code == 0 || misc[MR_FAIL + c] < misc[MR_FAIL + code])) code = c;
  I.trap_code = code;
  I.fit_check_failed = misc[MR_FAIL] < limit ? 1 : 0;
  if (info) *info = I;
  if (code) return fail(CPL_ERR_TRAP, "a cell hit a trap condition (code " + std::to_string(code) + ")");
  if (I.fit_check_failed) return fail(CPL_ERR_NOFIT, "circle possibly does not fit (middle/end fit check)");
  return CPL_OK;
}

// The printed line of a record, in the reference's printf formats
// (start_configuration.cu:132,147,158,207,289; middle_configuration.cu:182,193;
// end_configuration.cu:124,135,245,256).
int format_record(int family, const double* r, char* line, size_t cap) {
  const int kind = (int)r[0];
  const char* head = kind == 3 ? "Start configuration: Circle possibly does not fit - "
                               : kCritText[family][kind == 2 ? 2 : 1];
  if (family == 1 || family == 4)
    return snprintf(line, cap, "%sla = [%.19g,%.19g], r1 = [%.19g,%.19g], r2 = [%.19g,%.19g]\n", head, r[1], r[2],
                    r[3], r[4], r[5], r[6]);
  return snprintf(line, cap, "%sla = [%.19g,%.19g], r1 = [%.19g,%.19g], r2 = [%.19g,%.19g], r3 = [%.19g,%.19g]\n",
                  head, r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]);
}

// Appends the lines of `have` records to out_text (as many as fit) and
// completes the report's line counts and text size.
void write_lines(int family, const double* recs, uint64_t have, char* out_text, size_t out_text_cap,
                 cpl_mode_r_info& I) {
  size_t used = 0, needed = 1;
  if (out_text && out_text_cap) out_text[0] = 0;
  for (uint64_t k = 0; k < have; ++k) {
    char line[512];
    const int len = format_record(family, recs + 9 * k, line, sizeof line);
    if (len < 0) continue;
    needed += (size_t)len;
    if (out_text && used + (size_t)len + 1 <= out_text_cap) {
      std::memcpy(out_text + used, line, (size_t)len + 1);
      used += (size_t)len;
      ++I.lines_written;
    }
  }
  I.text_bytes_needed = needed;
}

}  // namespace

int cpl_first_problematic_ex(int family, cpl_iv outer_la, const int32_t* subdiv, uint32_t flags,
                             double* first_problematic_out, cpl_mode_r_info* info, char* out_text,
                             size_t out_text_cap) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!first_problematic_out) return fail(CPL_ERR_ARG, "cpl_first_problematic: null output");
  if (flags & ~(uint32_t)(CPL_MR_OUTPUT_CRITICALS | CPL_MR_EXHAUSTIVE))
    return fail(CPL_ERR_ARG, "cpl_first_problematic_ex: unknown flag");
  int32_t n[4];
  if ((rc = mode_r_subdiv(family, subdiv, n))) return rc;
  const bool out = (flags & CPL_MR_OUTPUT_CRITICALS) != 0;
  const uint32_t crit_cap = out ? CPL_MR_MAX_RECORDS : 0u;
  if ((rc = g.mr_ranges.ensure((size_t)n[0] * 4 * sizeof(double))) || (rc = g.mr_misc.ensure(MODE_R_MISC_BYTES)))
    return rc;
  if (crit_cap && (rc = g.mr_crit.ensure((size_t)crit_cap * 9 * sizeof(double)))) return rc;
  if ((rc = mode_r_enqueue(family, outer_la, n, out, (flags & CPL_MR_EXHAUSTIVE) != 0, g.mr_ranges.as<double>(),
                           g.mr_misc.as<unsigned long long>(), crit_cap ? g.mr_crit.as<double>() : nullptr, crit_cap)))
    return rc;
  std::vector<unsigned long long> misc(MODE_R_MISC_BYTES / 8);
  HIP_TRY(hipMemcpyAsync(misc.data(), g.mr_misc.p, MODE_R_MISC_BYTES, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  cpl_mode_r_info I;
  const int result = mode_r_finish(out, misc.data(), first_problematic_out, &I);
  const std::string err = g_err;
  if (out) {
    const uint64_t have = std::min<uint64_t>(I.lines_found, crit_cap);
    std::vector<double> recs((size_t)have * 9);
    if (have) HIP_TRY(hipMemcpy(recs.data(), g.mr_crit.p, recs.size() * sizeof(double), hipMemcpyDeviceToHost));
    write_lines(family, recs.data(), have, out_text, out_text_cap, I);
  }
  if (info) *info = I;
  g_err = err;
  return result;
}

// Host staging of the no-allocation entry point (static: the entry point
// allocates nothing; calls are not reentrant, as every entry point here).
static unsigned long long g_noalloc_misc[MODE_R_MISC_BYTES / 8];
static double g_noalloc_recs[(size_t)CPL_NOALLOC_MAX_RECORDS * 9];

int cpl_first_problematic_noalloc(int family, cpl_iv outer_la, const int32_t* subdiv, uint32_t flags,
                                  double* first_problematic_out, cpl_mode_r_info* info, char* out_text,
                                  size_t out_text_cap) {
  int rc = ensure_init();
  if (rc) return rc;
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != g.device)
    return fail(CPL_ERR_ARG, "cpl_first_problematic_noalloc: the library context belongs to device " +
                                 std::to_string(g.device) + ", the current device is " + std::to_string(dev));
  if (!first_problematic_out) return fail(CPL_ERR_ARG, "cpl_first_problematic_noalloc: null output");
  if (flags & ~(uint32_t)(CPL_MR_OUTPUT_CRITICALS | CPL_MR_EXHAUSTIVE))
    return fail(CPL_ERR_ARG, "cpl_first_problematic_noalloc: unknown flag");
  int32_t n[4];
  if ((rc = mode_r_subdiv(family, subdiv, n))) return rc;
  if (n[0] > (int32_t)CPL_NOALLOC_MAX_LA)
    return fail(CPL_ERR_ARG, "cpl_first_problematic_noalloc: " + std::to_string(n[0]) +
                                 " lambda subdivisions exceed the scratch's maximum of " +
                                 std::to_string(CPL_NOALLOC_MAX_LA));
  const bool out = (flags & CPL_MR_OUTPUT_CRITICALS) != 0;
  double* ranges = reinterpret_cast<double*>(g.noalloc_scratch);
  unsigned long long* misc = reinterpret_cast<unsigned long long*>(g.noalloc_scratch + kNoAllocRangesBytes);
  double* crit = reinterpret_cast<double*>(g.noalloc_scratch + kNoAllocRangesBytes + MODE_R_MISC_BYTES);
  if ((rc = mode_r_enqueue(family, outer_la, n, out, (flags & CPL_MR_EXHAUSTIVE) != 0, ranges, misc,
                           out ? crit : nullptr, CPL_NOALLOC_MAX_RECORDS)))
    return rc;
  HIP_TRY(hipMemcpyAsync(g_noalloc_misc, misc, MODE_R_MISC_BYTES, hipMemcpyDeviceToHost, g.stream));
  HIP_TRY(hipStreamSynchronize(g.stream));
  cpl_mode_r_info I;
  const int result = mode_r_finish(out, g_noalloc_misc, first_problematic_out, &I);
  const std::string err = g_err;
  if (out) {
    const uint64_t have = std::min<uint64_t>(I.lines_found, CPL_NOALLOC_MAX_RECORDS);
    if (have) {
      HIP_TRY(hipMemcpyAsync(g_noalloc_recs, crit, have * 9 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
      HIP_TRY(hipStreamSynchronize(g.stream));
    }
    write_lines(family, g_noalloc_recs, have, out_text, out_text_cap, I);
  }
  if (info) *info = I;
  g_err = err;
  return result;
}

int cpl_first_problematic(int family, cpl_iv outer_la, const int32_t* subdiv, int output_criticals,
                          double* first_problematic_out, uint64_t* cells_evaluated_out, char* out_text,
                          size_t out_text_cap) {
  cpl_mode_r_info info;
  std::memset(&info, 0, sizeof info);
  const int rc = cpl_first_problematic_ex(family, outer_la, subdiv, output_criticals ? CPL_MR_OUTPUT_CRITICALS : 0u,
                                          first_problematic_out, &info, out_text, out_text_cap);
  if (cells_evaluated_out) *cells_evaluated_out = info.cells_evaluated;
  return rc;
}

}  // extern "C"
