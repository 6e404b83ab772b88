// C entry points of the CPU oracle (liboracle_cpl.so) — TEST INFRASTRUCTURE.
// Used by tests/ through ctypes and by the CPU-baseline leg of bench.py; the
// product never links it.  Functions mirror the product's C ABI (include/cpl.h)
// so that tests compare the two on identical inputs.
#include "cpl_oracle.hpp"
#include "../include/cpl.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

using namespace cplo;

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t bits(double d) {
  uint64_t b;
  std::memcpy(&b, &d, 8);
  return b;
}

struct Item {
  double e[8];  // la.lb la.ub r1.lb r1.ub r2.lb r2.ub p3.lb p3.ub
  uint32_t root;
  uint32_t la_idx;   // index of the la cell at its own depth (lbrsp table key)
  uint8_t family;
  uint8_t depth;
  uint16_t splits;   // 4 bits per axis: number of bisections so far
};

// Order-independent worklist digest: identical definition in the HIP product.
inline uint64_t item_hash(const Item& it) {
  uint64_t h = 0x243F6A8885A308D3ull;
  for (int k = 0; k < 8; ++k) h = mix64(h ^ bits(it.e[k]));
  uint64_t meta = (uint64_t(it.root) << 32) | (uint64_t(it.family) << 24) | (uint64_t(it.depth) << 16) | it.splits;
  return mix64(h ^ meta);
}

int axis_splits(uint16_t s, int a) { return (s >> (4 * a)) & 15; }

struct RunState {
  const cpl_refine_cfg* cfg;
  std::vector<int> fams;
  // lbrsp memo keyed by la.lb bits (values are deterministic functions of la.lb)
  std::mutex memo_mu;
  std::map<uint64_t, double> memo;
};

double lbrsp_memo(RunState& st, double la_lb) {
  uint64_t key = bits(la_lb);
  {
    std::lock_guard<std::mutex> g(st.memo_mu);
    auto it = st.memo.find(key);
    if (it != st.memo.end()) return it->second;
  }
  double v = lower_bound_radius_single_placement(la_lb);
  std::lock_guard<std::mutex> g(st.memo_mu);
  st.memo[key] = v;
  return v;
}

struct LocalStats {
  uint64_t evaluations[CPL_MAX_ROUNDS] = {};
  uint64_t verdicts[CPL_MAX_ROUNDS][8] = {};
  uint64_t fam_verdicts[CPL_MAX_ROUNDS][NUM_FAMILIES][8] = {};
  uint64_t flops[NUM_FAMILIES][8] = {};   // counting build only: flops by family and verdict
  uint64_t digest[CPL_MAX_ROUNDS] = {};
  uint64_t vdigest[CPL_MAX_ROUNDS] = {};
  uint64_t certified = 0, unresolved = 0;
};

// Axis chosen for a bisection in split mode 0: the widest eligible axis in
// units of its root extent, i.e. the smallest grid[a] << splits[a]; ties go to
// the lowest axis.
int widest_axis(const cpl_refine_cfg* cfg, uint16_t splits) {
  int best = -1;
  uint64_t bw = 0;
  for (int a = 0; a < CPL_NUM_AXES; ++a) {
    if (!(cfg->split_axes & (1u << a))) continue;
    uint64_t w = uint64_t(cfg->grid[a]) << axis_splits(splits, a);
    if (best < 0 || w < bw) { best = a; bw = w; }
  }
  return best;
}

Item child_of(const Item& p, int axis, int c) {
  Item ch = p;
  IV iv(p.e[2 * axis], p.e[2 * axis + 1]);
  IV s = get_subinterval(iv, c, 2);
  ch.e[2 * axis] = s.lb;
  ch.e[2 * axis + 1] = s.ub;
  ch.splits = uint16_t(p.splits + (1u << (4 * axis)));
  ch.depth = uint8_t(p.depth + 1);
  if (axis == CPL_AXIS_LA) ch.la_idx = p.la_idx * 2 + c;
  return ch;
}

void process(RunState& st, LocalStats& ls, const Item& it, uint8_t* flags_out) {
  const cpl_refine_cfg* cfg = st.cfg;
  IV la(it.e[0], it.e[1]), r1(it.e[2], it.e[3]), r2(it.e[4], it.e[5]), p3(it.e[6], it.e[7]);
  double lbrsp = 0.0;
  if (cfg->clip) {
    uint64_t saved = flop_count();
    lbrsp = lbrsp_memo(st, la.lb);  // a table lookup on the device: not counted
    flop_count() = saved;
  }
  uint64_t f0 = flop_count();
  // CPL_OPT_LEARNED_ORDER: the oracle's stand-in for a proposal that always
  // votes for the second check (refinement rounds only; the root round has no
  // proposal)
  const bool second_first = (cfg->options & CPL_OPT_LEARNED_ORDER) && it.depth >= 1;
  uint8_t vb = eval_synthetic(it.family, la, r1, r2, p3, cfg->clip != 0, lbrsp, second_first);
  int v = vb & 7;
  int d = it.depth;
  ls.flops[it.family][v] += flop_count() - f0;
  ls.fam_verdicts[d][it.family][v]++;
  uint64_t h = item_hash(it);
  ls.evaluations[d]++;
  ls.verdicts[d][v]++;
  ls.digest[d] += h;
  ls.vdigest[d] += mix64(h ^ vb);
  if (vb & 8) *flags_out |= CPL_RF_NOFIT_FLAG;
  if (v == CRITICAL && d < cfg->max_depth) {
    if (cfg->split_mode == 0) {
      int a = widest_axis(cfg, it.splits);
      for (int c = 0; c < 2; ++c) process(st, ls, child_of(it, a, c), flags_out);
    } else {
      std::vector<Item> cur{it}, nxt;
      for (int a = 0; a < CPL_NUM_AXES; ++a) {
        if (!(cfg->split_axes & (1u << a))) continue;
        nxt.clear();
        for (const Item& x : cur)
          for (int c = 0; c < 2; ++c) {
            Item ch = child_of(x, a, c);
            ch.depth = uint8_t(it.depth + 1);
            nxt.push_back(ch);
          }
        cur.swap(nxt);
      }
      for (const Item& ch : cur) process(st, ls, ch, flags_out);
    }
    return;
  }
  switch (v) {
    case PROVEN: *flags_out |= CPL_RF_PROVEN; ls.certified++; break;
    case EMPTY: *flags_out |= CPL_RF_EMPTY; ls.certified++; break;
    case OOD: *flags_out |= CPL_RF_OOD; ls.certified++; break;
    case CRITICAL: *flags_out |= CPL_RF_CRITICAL; ls.unresolved++; break;
    case NOFIT: *flags_out |= CPL_RF_NOFIT; ls.unresolved++; break;
    default: *flags_out |= CPL_RF_TRAP; ls.unresolved++; break;
  }
}

Item root_item(const cpl_refine_cfg* cfg, uint64_t root, int family) {
  Item it{};
  uint64_t r = root;
  int idx[CPL_NUM_AXES];
  for (int a = CPL_NUM_AXES - 1; a >= 0; --a) {
    idx[a] = int(r % uint64_t(cfg->grid[a]));
    r /= uint64_t(cfg->grid[a]);
  }
  for (int a = 0; a < CPL_NUM_AXES; ++a) {
    IV s = get_subinterval(IV(cfg->box_lb[a], cfg->box_ub[a]), idx[a], cfg->grid[a]);
    it.e[2 * a] = s.lb;
    it.e[2 * a + 1] = s.ub;
  }
  it.root = uint32_t(root);
  it.la_idx = uint32_t(idx[0]);
  it.family = uint8_t(family);
  it.depth = 0;
  it.splits = 0;
  return it;
}

int hw_threads(int requested) {
  if (requested > 0) return requested;
  unsigned n = std::thread::hardware_concurrency();
  return n ? int(n) : 1;
}

}  // namespace

extern "C" {

int cplo_counting_build(void) { return CPLO_COUNTING; }

int cplo_prim(int op, const double* a, const double* b, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
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
      case 15: r = std::sqrt(x); break;
      case 16: r = add_rd32((float)x, (float)y); break;
      case 17: r = add_ru32((float)x, (float)y); break;
      case 18: r = sub_rd32((float)x, (float)y); break;
      case 19: r = sub_ru32((float)x, (float)y); break;
      case 20: r = mul_rd32((float)x, (float)y); break;
      case 21: r = mul_ru32((float)x, (float)y); break;
      case 22: r = div_rd32((float)x, (float)y); break;
      case 23: r = div_ru32((float)x, (float)y); break;
      case 27: r = f32_rd(x); break;
      case 28: r = f32_ru(x); break;
      case 29: r = sin_rn(x); break;
      case 30: r = cos_rn(x); break;
      case 31: r = acos_rn(x); break;
      case 32: r = dsin_rd(x); break;
      case 33: r = dsin_ru(x); break;
      case 34: r = dcos_rd(x); break;
      case 35: r = dcos_ru(x); break;
      case 36: r = dacos_rd(x); break;
      case 37: r = dacos_ru(x); break;
      case 38: {  // Atomic_ordered_double::transform, the u64 carried in the double's bits
        uint64_t o = ordered_from_double(x);
        std::memcpy(&r, &o, 8);
        break;
      }
      case 39: {  // inverse transform of the u64 carried in x's bits
        uint64_t o;
        std::memcpy(&o, &x, 8);
        r = double_from_ordered(o);
        break;
      }
      default: return CPL_ERR_ARG;
    }
    out[i] = r;
  }
  return CPL_OK;
}

// Elementwise interval operations (same op codes as cpl_interval_op).
int cplo_interval_op(int op, const double* a, const double* b, double* out, int* trap, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const IV x(a[2 * i], a[2 * i + 1]);
    const IV y = b ? IV(b[2 * i], b[2 * i + 1]) : IV(0.0, 0.0);
    IV r(0.0, 0.0);
    int t = 0;
    try {
      switch (op) {
        case 0: r = x + y; break;
        case 1: r = x - y; break;
        case 2: r = x * y; break;
        case 3: r = x / y; break;
        case 4: r = sqrt(x); break;
        case 5: r = x.square(); break;
        case 6: r = x.abs(); break;
        case 7: r = reciprocal(x); break;
        case 8: r = x / y.lb; break;
        case 9: r = IV(x.lb, x.lb) / y; break;
        case 10: r = sin(x); break;
        case 11: r = cos(x); break;
        case 12: r = asin(x); break;
        case 13: r = acos(x); break;
        case 14: r = x * y.lb; break;
        case 15: r = -x; break;
        default: return CPL_ERR_ARG;
      }
    } catch (const Trap& e) {
      t = e.code;
      r = IV(0.0, 0.0);
    }
    out[2 * i] = r.lb;
    out[2 * i + 1] = r.ub;
    trap[i] = t;
  }
  return CPL_OK;
}

// Geometry functions on intervals, for containment tests.  in: n records of
// 6 doubles (la.lb, la.ub, ra.lb, ra.ub, rb.lb, rb.ub); out: n x (lb, ub);
// trap: n ints (0 or trap code).
//   fn 0 inner_inner  1 outer_inner  2 inner_outer  3 outer_outer(ra, rb)
//   fn 4 semicircle_outer(ra)  5 semicircle_inner(la, ra)
//   fn 6 compute_collision_angle(lc = la, sc = ra, ab = rb)
int cplo_geom(int fn, const double* in, double* out, int* trap, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const double* p = in + 6 * i;
    IV la(p[0], p[1]), ra(p[2], p[3]), rb(p[4], p[5]);
    IV r(0, 0);
    int t = 0;
    try {
      switch (fn) {
        case 0: r = collision_angle_inner_inner(la, ra, rb); break;
        case 1: r = collision_angle_outer_inner(la, ra, rb); break;
        case 2: r = collision_angle_inner_outer(la, ra, rb); break;
        case 3: r = collision_angle_outer_outer(ra, rb); break;
        case 4: r = semicircle_angle_outer(ra); break;
        case 5: r = semicircle_angle_inner(la, ra); break;
        case 6: r = compute_collision_angle(la, ra, rb); break;
        default: return CPL_ERR_ARG;
      }
    } catch (const Trap& tr) {
      t = tr.code;
    }
    out[2 * i] = r.lb;
    out[2 * i + 1] = r.ub;
    if (trap) trap[i] = t;
  }
  return CPL_OK;
}

// Admissible ranges of family f (a checker for the range functions): in: n
// records of 6 doubles (la, r1, r2 intervals); out: n x 6 doubles =
// compute_r1_range(la), r2 range(la, r1), r3 range(la, r1, r2) (the r3 range of
// the 3-d families; zeros for the 2-d ones).
int cplo_ranges(int family, const double* in, double* out, size_t n) {
  if (family < 0 || family >= NUM_FAMILIES) return CPL_ERR_ARG;
  for (size_t i = 0; i < n; ++i) {
    const double* p = in + 6 * i;
    IV la(p[0], p[1]), r1(p[2], p[3]), r2(p[4], p[5]);
    IV a = compute_r1_range(la, lower_bound_radius_single_placement(la.lb));
    IV b = r2_range_for(family, la, r1);
    IV c = family_is_3d(family) ? r3_range_for(family, la, r1, r2) : IV(0.0, 0.0);
    double* o = out + 6 * i;
    o[0] = a.lb; o[1] = a.ub; o[2] = b.lb; o[3] = b.ub; o[4] = c.lb; o[5] = c.ub;
  }
  return CPL_OK;
}

// The area comparisons of the output_criticals evaluation of n cells (8
// doubles each, as cplo_eval_kernel_body): out n x 16 doubles = up to 8 pairs
// (lower bound of the disk area, upper bound of critical_ratio x used area) in
// evaluation order; count n ints = pairs recorded, or -1 when the cell is
// decided before the criteria or traps.
int cplo_area_terms(int family, const double* cells, size_t n, double* out, int* count) {
  if (family < 0 || family >= NUM_FAMILIES) return CPL_ERR_ARG;
  AreaProbe& p = area_probe();
  for (size_t i = 0; i < n; ++i) {
    const double* c = cells + 8 * i;
    p.on = 1;
    p.n = 0;
    int nofit = 0, early = -1, k = 0;
    try {
      eval_kernel_body_all(family, IV(c[0], c[1]), IV(c[2], c[3]), IV(c[4], c[5]), IV(c[6], c[7]), &nofit, &early);
      k = early >= 0 ? -1 : p.n;
    } catch (const Trap&) {
      k = -1;
    }
    p.on = 0;
    for (int j = 0; j < 16; ++j) out[16 * i + j] = j < 2 * p.n ? p.v[j] : 0.0;
    count[i] = k;
  }
  return CPL_OK;
}

int cplo_eval_kernel_body(int family, const double* cells, size_t n, uint8_t* verdicts) {
  if (family < 0 || family >= NUM_FAMILIES) return CPL_ERR_ARG;
  for (size_t i = 0; i < n; ++i) {
    const double* c = cells + 8 * i;
    verdicts[i] = eval_kernel_body(family, IV(c[0], c[1]), IV(c[2], c[3]), IV(c[4], c[5]), IV(c[6], c[7]));
  }
  return CPL_OK;
}

int cplo_eval_cells(int family, int clip, const double* const soa[8], size_t n, uint8_t* verdicts, int nthreads) {
  if (family < 0 || family >= NUM_FAMILIES) return CPL_ERR_ARG;
  int T = hw_threads(nthreads);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t s = next.fetch_add(1024);
      if (s >= n) break;
      size_t e = std::min(n, s + 1024);
      for (size_t i = s; i < e; ++i) {
        IV la(soa[0][i], soa[1][i]);
        double lbrsp = clip ? lower_bound_radius_single_placement(la.lb) : 0.0;
        verdicts[i] = eval_synthetic(family, la, IV(soa[2][i], soa[3][i]), IV(soa[4][i], soa[5][i]),
                                     IV(soa[6][i], soa[7][i]), clip != 0, lbrsp);
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  return CPL_OK;
}

// Synthetic refinement over roots [root_begin, root_end) (all roots when
// root_end == 0).  root_flags: (root_end - root_begin) * nfam bytes.
// flops_out (may be NULL): 6 x 8 sums of counted FP64 operations by family
// and verdict (nonzero only in the counting build, liboracle_cpl_count.so).
int cplo_refine_run(const cpl_refine_cfg* cfg, cpl_refine_stats* stats, uint8_t* root_flags,
                    uint64_t root_begin, uint64_t root_end, uint64_t root_stride, int nthreads,
                    uint64_t* flops_out) {
  if (!cfg || !stats) return CPL_ERR_ARG;
  if (cfg->max_depth < 0 || cfg->max_depth >= CPL_MAX_ROUNDS) return CPL_ERR_ARG;
  RunState st;
  st.cfg = cfg;
  for (int f = 0; f < NUM_FAMILIES; ++f)
    if (cfg->family_mask & (1u << f)) st.fams.push_back(f);
  uint64_t n_roots = 1;
  for (int a = 0; a < CPL_NUM_AXES; ++a) n_roots *= uint64_t(cfg->grid[a]);
  if (root_end == 0 || root_end > n_roots) root_end = n_roots;
  if (root_stride == 0) root_stride = 1;
  uint64_t count = root_end > root_begin ? (root_end - root_begin + root_stride - 1) / root_stride : 0;
  const size_t nf = st.fams.size();
  int T = hw_threads(nthreads);
  std::vector<LocalStats> locals(T);
  std::atomic<uint64_t> next{0};
  auto work = [&](int tid) {
    LocalStats& ls = locals[tid];
    for (;;) {
      uint64_t s = next.fetch_add(256);
      if (s >= count) break;
      uint64_t e = std::min(count, s + 256);
      for (uint64_t k = s; k < e; ++k) {
        uint64_t root = root_begin + k * root_stride;
        for (size_t fi = 0; fi < nf; ++fi) {
          uint8_t fl = 0;
          process(st, ls, root_item(cfg, root, st.fams[fi]), &fl);
          if (root_flags) root_flags[k * nf + fi] = fl;
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& t : th) t.join();
  std::memset(stats, 0, sizeof(*stats));
  for (const LocalStats& ls : locals) {
    for (int d = 0; d < CPL_MAX_ROUNDS; ++d) {
      stats->evaluations[d] += ls.evaluations[d];
      for (int v = 0; v < 8; ++v) stats->verdicts[d][v] += ls.verdicts[d][v];
      for (int f = 0; f < NUM_FAMILIES; ++f)
        for (int v = 0; v < 8; ++v) stats->family_verdicts[d][f][v] += ls.fam_verdicts[d][f][v];
      stats->digest[d] += ls.digest[d];
      stats->verdict_digest[d] += ls.vdigest[d];
    }
    stats->certified += ls.certified;
    stats->unresolved += ls.unresolved;
  }
  if (flops_out)
    for (const LocalStats& ls : locals)
      for (int f = 0; f < NUM_FAMILIES; ++f)
        for (int v = 0; v < 8; ++v) flops_out[f * 8 + v] += ls.flops[f][v];
  int rounds = 0;
  for (int d = 0; d < CPL_MAX_ROUNDS; ++d) {
    stats->total_evaluations += stats->evaluations[d];
    if (stats->evaluations[d]) rounds = d + 1;
  }
  stats->rounds = rounds;
  return CPL_OK;
}

// Reference-grid search (Mode R) for one lambda interval, evaluating every
// cell of the grid (start_configuration.cu:114-167, :190-215, :267-298;
// middle_configuration.cu:156-205; end_configuration.cu:103-146, :212-268).
//   first: min la.lb over cells with a failing criterion (DBL_MAX if none).
//   A trap, or the fit check of the middle and end families, is reported only
//   where the reference is bound to reach it: every thread of the reference
//   walks its cells in increasing la and returns at its first failing cell, so
//   every cell with la.lb < first is reached, while cells at or above first may
//   or may not be (that depends on the thread mapping).  rc = CPL_ERR_TRAP when
//   a trapping cell has la.lb < first (trap_code: the code whose smallest
//   trapping la.lb is smallest, then the smallest code), else CPL_ERR_NOFIT
//   when a fit-check cell has la.lb < first.  With output_criticals the
//   reference evaluates every cell and every criterion: every trap and fit
//   failure counts, and recs receives one record per printed line (9 doubles:
//   kind, la, r1, r2, r3 as printed; kind 1/2 = failing criterion 1/2, 3 = the
//   start family's "possibly does not fit" line), sorted, at most rec_cap of
//   the n_recs_out found.  nofit_flags_out: cells of the start family whose
//   fit check fired (the reference prints a line for each it reaches).
int cplo_first_problematic(int family, double outer_lb, double outer_ub, const int32_t* subdiv,
                           int output_criticals, double* first_out, uint64_t* evaluated_out, int* trap_code_out,
                           double* recs, uint64_t rec_cap, uint64_t* n_recs_out, uint64_t* nofit_flags_out,
                           int nthreads) {
  if (family < 0 || family >= NUM_FAMILIES || !subdiv || !first_out) return CPL_ERR_ARG;
  const int nla = subdiv[0], nr1 = subdiv[1], nr2 = subdiv[2];
  const int nr3 = family_is_3d(family) ? subdiv[3] : 1;
  if (nla < 1 || nr1 < 1 || nr2 < 1 || nr3 < 1) return CPL_ERR_ARG;
  IV outer(outer_lb, outer_ub);
  int T = hw_threads(nthreads);
  std::atomic<int> next{0};
  std::mutex mu;
  uint64_t best = ordered_from_double(DBL_MAX);
  uint64_t fail_min[6];  // [0] fit check, [c] trap code c
  for (uint64_t& v : fail_min) v = ~0ull;
  uint64_t evaluated = 0, nofit_flags = 0;
  std::vector<std::array<double, 9>> lines;
  auto work = [&]() {
    uint64_t ev = 0, nf = 0, b = ordered_from_double(DBL_MAX);
    uint64_t fm[6];
    for (uint64_t& v : fm) v = ~0ull;
    std::vector<std::array<double, 9>> mine;
    for (;;) {
      int job = next.fetch_add(1);
      if (job >= nla * nr1) break;
      int ia = job / nr1, i1 = job % nr1;
      IV la = get_subinterval(outer, ia, nla);
      const uint64_t o = ordered_from_double(la.lb);
      IV r1_range = compute_r1_range(la, lower_bound_radius_single_placement(la.lb));
      IV r1_ = get_subinterval(r1_range, i1, nr1);
      IV r2_range = r2_range_for(family, la, r1_);
      for (int i2 = 0; i2 < nr2; ++i2) {
        IV r2_ = get_subinterval(r2_range, i2, nr2);
        IV r3_range(0, 0);
        if (family_is_3d(family)) {
          r3_range = r3_range_for(family, la, r1_, r2_);
          if (family == END_3D && r3_range.empty()) continue;  // end_configuration.cu:222-224
        }
        for (int i3 = 0; i3 < nr3; ++i3) {
          IV r3 = family_is_3d(family) ? get_subinterval(r3_range, i3, nr3) : IV(0, 0);
          ++ev;
          if (!output_criticals) {
            uint8_t vb = eval_kernel_body(family, la, r1_, r2_, r3);
            int v = vb & 7;
            if (vb & 8) ++nf;
            if (v == CRITICAL) b = std::min(b, o);
            else if (v == NOFIT) fm[0] = std::min(fm[0], o);
            else if (v == TRAP) fm[vb >> 4] = std::min(fm[vb >> 4], o);
            continue;
          }
          int nofit_flag = 0, early = -1, mask = 0;
          try {
            mask = eval_kernel_body_all(family, la, r1_, r2_, r3, &nofit_flag, &early);
          } catch (const Trap& t) {
            fm[t.code] = std::min(fm[t.code], o);
          }
          if (early == NOFIT) fm[0] = std::min(fm[0], o);
          if (early >= 0) continue;
          IV r2 = r2_;
          if (family_is_3d(family)) r2.tighten_lb(r3.lb);
          IV r1 = r1_;
          r1.tighten_lb(r2.lb);
          auto line = [&](int kind) {
            mine.push_back({double(kind), la.lb, la.ub, r1.lb, r1.ub, r2.lb, r2.ub, r3.lb, r3.ub});
          };
          if (nofit_flag) {
            ++nf;
            line(3);
          }
          if (mask) b = std::min(b, o);
          for (int k = 0; k < 2; ++k)
            if (mask & (1 << k)) line(k + 1);
        }
      }
    }
    std::lock_guard<std::mutex> g(mu);
    evaluated += ev;
    nofit_flags += nf;
    best = std::min(best, b);
    for (int k = 0; k < 6; ++k) fail_min[k] = std::min(fail_min[k], fm[k]);
    lines.insert(lines.end(), mine.begin(), mine.end());
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  *first_out = double_from_ordered(best);
  if (evaluated_out) *evaluated_out = evaluated;
  if (nofit_flags_out) *nofit_flags_out = nofit_flags;
  std::sort(lines.begin(), lines.end());
  if (n_recs_out) *n_recs_out = lines.size();
  if (recs)
    for (size_t k = 0; k < lines.size() && k < rec_cap; ++k) std::memcpy(recs + 9 * k, lines[k].data(), 9 * 8);
  const uint64_t threshold = output_criticals ? ~0ull : best;
  int code = 0;
  for (int c = 1; c < 6; ++c)
    if (fail_min[c] < threshold && (code == 0 || fail_min[c] < fail_min[code])) code = c;
  if (trap_code_out) *trap_code_out = code;
  if (code) return CPL_ERR_TRAP;
  if (fail_min[0] < threshold) return CPL_ERR_NOFIT;
  return CPL_OK;
}

}  // extern "C"
