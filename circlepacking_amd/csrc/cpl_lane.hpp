// Lane geometry, criteria and the per-cell kernel body on the device.
//
// Follows lane_packing/collision_angles.cu, lower_bound_radius_single_placement.cu
// and the three *_configuration.cu files of the reference formula for formula
// (each function cites its source lines).  Differences in mechanism only:
//  * trap() becomes a trap code in `trap` (first one wins; callers stop at once);
//  * the kernel body returns a verdict byte instead of calling store_min/printf;
//  * geometry and criteria are templates on the scalar type: T = double is the
//    certifying evaluation, T = float the outward-rounded prefilter (its
//    constants are rounded outward so every FP32 bound encloses the FP64 one).
#pragma once

#include <type_traits>

#include "cpl_asin.hpp"

namespace cpl {

enum : int { F_START = 0, F_START_END_2D = 1, F_START_END_3D = 2, F_MIDDLE = 3, F_END_2D = 4, F_END_3D = 5 };
enum : int { V_PROVEN = 0, V_CRITICAL = 1, V_NOFIT = 2, V_EMPTY = 3, V_OOD = 4, V_TRAP = 5 };
enum : int { T_ASIN = 1, T_OBT_X = 2, T_OBT_Q = 3, T_ACU_X = 4 };

__device__ __forceinline__ constexpr bool is_3d(int f) {
  return f == F_START || f == F_START_END_3D || f == F_MIDDLE || f == F_END_3D;
}
__device__ __forceinline__ uint8_t vbyte(int v, int nofit_flag, int detail) {
  return (uint8_t)(v | (nofit_flag ? 8 : 0) | (detail << 4));
}

constexpr double critical_ratio = 0.5642;   // values.hpp:28
constexpr double lb_pi = 3.141592653589793; // values.hpp:29

template <class T> struct LaneConst;
template <> struct LaneConst<double> {
  static constexpr double crit = critical_ratio;  // used as an upper bound (mul_ru)
  static constexpr double lbpi = lb_pi;           // used as a lower bound (mul_rd)
};
template <> struct LaneConst<float> {
  static constexpr float crit = 0x1.20ded4p-1f;   // 0.5642 rounded up
  static constexpr float lbpi = 0x1.921fb4p+1f;   // 3.141592653589793 rounded down
};

// Interval asin (interval.cuh:617-624).
template <class T>
__device__ __forceinline__ IVT<T> asin_iv(IVT<T> x, int& trap) {
  if (x.lb < T(-1) || x.ub > T(1)) {
    if (!trap) trap = T_ASIN;
    return IVT<T>(T(0), T(0));
  }
  return IVT<T>(dasin_rd(x.lb), dasin_ru(x.ub));
}

// Result of an out-of-line geometry function: the angle interval and the trap
// code (0: none), returned in registers (a trap flag passed by reference would
// live in memory).
template <class T>
struct Angle {
  IVT<T> v;
  int trap;
};

// collision_angles.cu:37-117 (with the obtuseness test of :29-31)
//
// W names the endpoints of the result the caller reads (bit 0: lb, bit 1: ub;
// every caller on the path reads one).  An arcsine that reaches only an unread
// endpoint is not evaluated; the quotients, their domain checks and therefore
// every trap decision are those of the full computation, and when both branches
// may apply the obtuse branch's result is computed in full, because its
// emptiness decides between taking and joining the acute one.  Read endpoints
// are bit-identical to the full computation's; unread ones are unspecified.
template <class T>
__device__ __forceinline__ bool asin_out_of_domain(IVT<T> x) {  // asin_iv's trap test
  return x.lb < T(-1) || x.ub > T(1);
}

// Single-endpoint quotients (FP64, W = 1 or 2).  For a >= 0 and 0 < o, both
// non-empty, finite and o within [2^-500, 2^500), the interval a / o is
// [RD(a.lb / o.ub), RU(a.ub / o.lb)] with lb <= ub, so its tests reduce to one
// endpoint and exact comparisons: after tighten_ub(1) the quotient is empty iff
// RD(a.lb / o.ub) > 1, i.e. a.lb >= succ(1) o.ub (the sign of one fma, exact in
// this range); its upper bound exceeds 1 iff a.ub > o.lb.  Where every lane of
// the wave meets these conditions the kernel computes only the quotient
// endpoints its read endpoint needs (two directed divisions instead of four).
__device__ __forceinline__ bool quot_fast_ok(IV a, IV o) {
  return a.lb >= 0.0 && a.lb <= a.ub && a.ub < 0x1p+500 && o.lb >= 0x1p-500 && o.lb <= o.ub && o.ub < 0x1p+500;
}
__device__ __forceinline__ bool rd_quot_above_one(double a, double b) {  // RD(a / b) > 1
  return __builtin_fma(-0x1.0000000000001p+0, b, a) >= 0.0;
}
template <class T, int W>
__device__ __attribute__((noinline)) Angle<T> compute_collision_angle_r(IVT<T> lc, IVT<T> sc, IVT<T> ab) {
  static_assert(W >= 1 && W <= 3, "endpoint selection");
  int trap = 0;
  using I = IVT<T>;
  lc.tighten_lb(sc.lb);
  sc.tighten_ub(lc.ub);
  I result(T(0), T(-1));
  I lc_sq, sc_sq;
  if constexpr (std::is_same<T, double>::value) {
    // both squares in one statement when square()'s own fast case holds for both
    if (all_lanes(lc.lb >= T(0) && lc.ub > T(0) && sc.lb >= T(0) && sc.ub > T(0))) {
      mul_pair2(lc.lb, lc.lb, lc.ub, lc.ub, sc.lb, sc.lb, sc.ub, sc.ub, lc_sq.lb, lc_sq.ub, sc_sq.lb, sc_sq.ub);
    } else {
      lc_sq = lc.square();
      sc_sq = sc.square();
    }
  } else {
    lc_sq = lc.square();
    sc_sq = sc.square();
  }
  // is_obtuse (collision_angles.cu:29-31) on the squares computed once: the
  // directed-rounding statements are not merged by the compiler
  UB obt = lc_sq > sc_sq + ab.square();
  I add_term = (lc_sq - sc_sq) / ab;
  add_term.tighten_lb(T(0));

  if (possibly(obt)) {
    I lx, sx;  // 1/2 (add_term + ab), 1/2 (add_term - ab)
    if constexpr (std::is_same<T, double>::value) {
      half_sum_diff(0.5, add_term.lb, add_term.ub, ab.lb, ab.ub, lx.lb, lx.ub, sx.lb, sx.ub);
    } else {
      lx = T(0.5) * (add_term + ab);
      sx = T(0.5) * (add_term - ab);
    }
    sx.tighten_lb(T(0));
    sx.tighten_ub(sc.ub);
    lx.tighten_ub(lc.ub);
    if (sx.empty() || lx.empty()) {
      if (definitely(obt)) { trap = T_OBT_X; return Angle<T>{result, trap}; }
    } else {
      bool done = false;
      if constexpr (std::is_same<T, double>::value && W != 3) {
        if (all_lanes(definitely(obt) && quot_fast_ok(lx, lc) && quot_fast_ok(sx, sc))) {
          // one endpoint of each quotient (quot_fast_ok); both lie in [0, 1], no domain trap
          T lq, sq;
          bool empty;
          if (W == 1) {
            div_pair2<false, true>(lx.lb, lc.ub, sx.ub, sc.lb, lq, sq);  // lquot.lb, squot.ub
            if (T(1) < sq) sq = T(1);                                     // squot tightened to 1
            empty = lq > T(1) || rd_quot_above_one(sx.lb, sc.ub);
          } else {
            div_pair2<true, false>(lx.ub, lc.lb, sx.lb, sc.ub, lq, sq);  // lquot.ub, squot.lb
            if (T(1) < lq) lq = T(1);                                     // lquot tightened to 1
            empty = rd_quot_above_one(lx.lb, lc.ub) || sq > T(1);
          }
          if (empty) { trap = T_OBT_Q; return Angle<T>{result, trap}; }
          T rl = T(0), ru = T(0);
          T al, as;  // the two widened arcsines, side by side
          if (W == 1) {
            dasin_pair<false, true>(lq, sq, al, as);
            rl = sub_rd(al, as);
          } else {
            dasin_pair<true, false>(lq, sq, al, as);
            ru = sub_ru(al, as);
          }
          result = I(rl, ru);
          result.tighten_lb(T(0));
          done = true;
        }
      }
      if (!done) {
        I lquot = lx / lc;
        lquot.tighten_ub(T(1));
        I squot = sx / sc;
        squot.tighten_ub(T(1));
        if (lquot.empty() || squot.empty()) {
          if (definitely(obt)) { trap = T_OBT_Q; return Angle<T>{result, trap}; }
        } else {
          if (asin_out_of_domain(lquot) || asin_out_of_domain(squot)) {
            trap = T_ASIN;
            return Angle<T>{result, trap};
          }
          // result = asin(lquot) - asin(squot), lower bound clamped at 0.  With
          // A = dasin_rd(lq.lb), B = dasin_ru(lq.ub), C = dasin_ru(sq.ub),
          // D = dasin_rd(sq.lb): the widened arcsines enclose the monotone arcsine,
          // so A <= B and D <= C, rd(A - C) <= ru(B - D), and after tighten_lb(0)
          // the result is empty iff ru(B - D) < 0.  When the acute branch follows
          // (obtuseness undecided), that emptiness decides between taking and
          // joining the acute result, so the endpoint no caller reads is needed
          // only for it: not at all for an upper-only call (its lower bound is left
          // at -inf, clamped to 0), and for a lower-only call only when
          // lq.ub < sq.lb (otherwise B >= asin(lq.ub) >= asin(sq.lb) >= D and
          // ru >= 0; its upper bound is left at +inf).  Read endpoints and every
          // emptiness decision are those of the full computation.
          T rl = T(0), ru = T(0);
          const T inf = T(__builtin_huge_val());
          if constexpr (W == 3) {
            rl = sub_rd(dasin_rd(lquot.lb), dasin_ru(squot.ub));
            ru = sub_ru(dasin_ru(lquot.ub), dasin_rd(squot.lb));
          } else if constexpr (W == 1) {
            rl = sub_rd(dasin_rd(lquot.lb), dasin_ru(squot.ub));
            if (!definitely(obt)) ru = lquot.ub < squot.lb ? sub_ru(dasin_ru(lquot.ub), dasin_rd(squot.lb)) : inf;
          } else {
            ru = sub_ru(dasin_ru(lquot.ub), dasin_rd(squot.lb));
            if (!definitely(obt)) rl = -inf;
          }
          result = I(rl, ru);
          result.tighten_lb(T(0));
        }
      }
    }
  }

  if (!definitely(obt)) {
    I lx, sx;  // 1/2 (ab + add_term), 1/2 (ab - add_term)
    if constexpr (std::is_same<T, double>::value) {
      half_sum_diff(0.5, ab.lb, ab.ub, add_term.lb, add_term.ub, lx.lb, lx.ub, sx.lb, sx.ub);
    } else {
      lx = T(0.5) * (ab + add_term);
      sx = T(0.5) * (ab - add_term);
    }
    sx.tighten_lb(T(0));
    if (sx.empty()) {
      if (!possibly(obt)) { trap = T_ACU_X; return Angle<T>{result, trap}; }
    } else {
      bool done = false;
      if constexpr (std::is_same<T, double>::value && W != 3) {
        if (all_lanes(quot_fast_ok(lx, lc) && quot_fast_ok(sx, sc))) {
          // one endpoint of each quotient; upper bounds above 1 trap as asin_iv does
          if (lx.ub > lc.lb) { trap = T_ASIN; return Angle<T>{result, trap}; }
          if (sx.ub > sc.lb) { trap = T_ASIN; return Angle<T>{result, trap}; }
          I r(T(0), T(0));
          T ql, qs, al, as;  // both quotients, then both widened arcsines, side by side
          if (W == 1) {
            div_pair2<false, false>(lx.lb, lc.ub, sx.lb, sc.ub, ql, qs);
            dasin_pair<false, false>(ql, qs, al, as);
            I beta(as, T(0));
            beta.tighten_lb(T(0));
            r.lb = add_rd(al, beta.lb);
          } else {
            div_pair2<true, true>(lx.ub, lc.lb, sx.ub, sc.lb, ql, qs);
            dasin_pair<true, true>(ql, qs, al, as);
            r.ub = add_ru(al, as);
          }
          if (result.empty()) result = r;
          else result.do_join(r);
          done = true;
        }
      }
      if (!done) {
        I lquot = lx / lc;
        if (asin_out_of_domain(lquot)) { trap = T_ASIN; return Angle<T>{result, trap}; }
        I squot = sx / sc;
        if (asin_out_of_domain(squot)) { trap = T_ASIN; return Angle<T>{result, trap}; }
        // r = asin(lquot) + asin(squot), the second's lower bound clamped at 0
        I r(T(0), T(0));
        if (W & 1) {
          I beta(dasin_rd(squot.lb), T(0));
          beta.tighten_lb(T(0));
          r.lb = add_rd(dasin_rd(lquot.lb), beta.lb);
        }
        if (W & 2) r.ub = add_ru(dasin_ru(lquot.ub), dasin_ru(squot.ub));
        if (result.empty()) result = r;
        else result.do_join(r);
      }
    }
  }
  return Angle<T>{result, trap};
}
template <int W, class T>
__device__ __forceinline__ IVT<T> compute_collision_angle(IVT<T> lc, IVT<T> sc, IVT<T> ab, int& trap) {
  const Angle<T> a = compute_collision_angle_r<T, W>(lc, sc, ab);
  if (a.trap) trap = a.trap;
  return a.v;
}

// collision_angles.cu:119-142
template <class T>
__device__ __attribute__((noinline)) Angle<T> collision_angle_inner_inner_r(IVT<T> la, IVT<T> ra, IVT<T> rb) {
  int trap = 0;
  using I = IVT<T>;
  T lb, ub;
  {
    I lc = la.lb + I(ra.ub, ra.ub);
    I sc = la.lb + I(rb.ub, rb.ub);
    I ab = I(ra.ub, ra.ub) + rb.ub;
    ub = compute_collision_angle<2>(lc, sc, ab, trap).ub;
    if (trap) return Angle<T>{I(T(0), T(0)), trap};
  }
  {
    I lc = la.ub + I(ra.lb, ra.lb);
    I sc = la.ub + I(rb.lb, rb.lb);
    I ab = I(ra.lb, ra.lb) + rb.lb;
    if (ab.lb <= T(0)) lb = T(0);
    else lb = compute_collision_angle<1>(lc, sc, ab, trap).lb;
  }
  return Angle<T>{I(lb, ub), trap};
}
template <class T>
__device__ __forceinline__ IVT<T> collision_angle_inner_inner(IVT<T> la, IVT<T> ra, IVT<T> rb, int& trap) {
  const Angle<T> a = collision_angle_inner_inner_r(la, ra, rb);
  if (a.trap) trap = a.trap;
  return a.v;
}

// collision_angles.cu:144-159 (outer_inner: ca = 1 - ra, cb = la + rb) and
// :176-191 (inner_outer: ca = la + ra, cb = 1 - rb).
// W: the endpoints the caller reads, as for compute_collision_angle_r (the
// join combines like endpoints only).
//
// The reference's three cases (ca definitely longer, cb definitely longer,
// either) are two calls at most per lane: the first with the operands in the
// order of the lane's case, the second, for the undecided lanes that did not
// trap, with the order swapped and joined.  A wave whose lanes fall in
// different cases runs the angle kernel at most twice instead of once per
// case; each lane's calls and their order are the reference's.
template <int W, class T>
__device__ __forceinline__ IVT<T> do_collision_angle(IVT<T> ca, IVT<T> cb, IVT<T> ab, int& trap) {
  UB ca_longer = (cb <= ca);
  const bool cb_first = definitely(!ca_longer);
  const bool either = !definitely(ca_longer) && !cb_first;
  IVT<T> result = compute_collision_angle<W>(cb_first ? cb : ca, cb_first ? ca : cb, ab, trap);
  if (either && !trap) result.do_join(compute_collision_angle<W>(cb, ca, ab, trap));
  return result;
}

// collision_angles.cu:161-174
template <class T>
__device__ __attribute__((noinline)) Angle<T> collision_angle_outer_inner_r(IVT<T> la, IVT<T> ra, IVT<T> rb) {
  int trap = 0;
  using I = IVT<T>;
  T lb, ub;
  if (possibly(T(2) * (ra + rb) <= T(1) - la)) {
    lb = T(0);
  } else {
    I pa(ra.lb, ra.lb), pb(rb.lb, rb.lb);
    lb = do_collision_angle<1>(T(1) - pa, la + pb, pa + pb, trap).lb;
    if (trap) return Angle<T>{I(T(0), T(0)), trap};
  }
  I pa(ra.ub, ra.ub), pb(rb.ub, rb.ub);
  ub = do_collision_angle<2>(T(1) - pa, la + pb, pa + pb, trap).ub;
  return Angle<T>{I(lb, ub), trap};
}
template <class T>
__device__ __forceinline__ IVT<T> collision_angle_outer_inner(IVT<T> la, IVT<T> ra, IVT<T> rb, int& trap) {
  const Angle<T> a = collision_angle_outer_inner_r(la, ra, rb);
  if (a.trap) trap = a.trap;
  return a.v;
}

// collision_angles.cu:193-205
template <class T>
__device__ __attribute__((noinline)) Angle<T> collision_angle_inner_outer_r(IVT<T> la, IVT<T> ra, IVT<T> rb) {
  int trap = 0;
  using I = IVT<T>;
  T lb, ub;
  if (possibly(T(2) * (ra + rb) <= T(1) - la)) {
    lb = T(0);
  } else {
    I pa(ra.lb, ra.lb), pb(rb.lb, rb.lb);
    lb = do_collision_angle<1>(la + pa, T(1) - pb, pa + pb, trap).lb;
    if (trap) return Angle<T>{I(T(0), T(0)), trap};
  }
  I pa(ra.ub, ra.ub), pb(rb.ub, rb.ub);
  ub = do_collision_angle<2>(la + pa, T(1) - pb, pa + pb, trap).ub;
  return Angle<T>{I(lb, ub), trap};
}
template <class T>
__device__ __forceinline__ IVT<T> collision_angle_inner_outer(IVT<T> la, IVT<T> ra, IVT<T> rb, int& trap) {
  const Angle<T> a = collision_angle_inner_outer_r(la, ra, rb);
  if (a.trap) trap = a.trap;
  return a.v;
}

// collision_angles.cu:207-231
template <class T>
__device__ __attribute__((noinline)) Angle<T> collision_angle_outer_outer_r(IVT<T> ra, IVT<T> rb) {
  int trap = 0;
  using I = IVT<T>;
  T lb, ub;
  {
    I lc = I(T(1), T(1)) - rb.ub;
    I sc = I(T(1), T(1)) - ra.ub;
    I ab = I(ra.ub, ra.ub) + rb.ub;
    ub = compute_collision_angle<2>(lc, sc, ab, trap).ub;
    if (trap) return Angle<T>{I(T(0), T(0)), trap};
  }
  {
    I lc = I(T(1), T(1)) - rb.lb;
    I sc = I(T(1), T(1)) - ra.lb;
    I ab = I(ra.lb, ra.lb) + rb.lb;
    if (ab.lb <= T(0)) lb = T(0);
    else lb = compute_collision_angle<1>(lc, sc, ab, trap).lb;
  }
  return Angle<T>{I(lb, ub), trap};
}
template <class T>
__device__ __forceinline__ IVT<T> collision_angle_outer_outer(IVT<T> ra, IVT<T> rb, int& trap) {
  const Angle<T> a = collision_angle_outer_outer_r(ra, rb);
  if (a.trap) trap = a.trap;
  return a.v;
}

// collision_angles.cu:233-238 (FP64: both quotients and both arcsines side by side)
template <class T>
__device__ __forceinline__ IVT<T> semicircle_angle_outer(IVT<T> r) {
  if constexpr (std::is_same<T, double>::value) {
    double ql, qu, yl, yu;
    div_pair2<false, true>(r.lb, add_ru(1.0, -r.lb), r.ub, add_rd(1.0, -r.ub), ql, qu);
    dasin_pair<false, true>(ql, qu, yl, yu);
    return IV(yl, yu);
  } else {
    return IVT<T>(dasin_rd(div_rd(r.lb, add_ru(T(1), -r.lb))), dasin_ru(div_ru(r.ub, add_rd(T(1), -r.ub))));
  }
}
// collision_angles.cu:240-246
template <class T>
__device__ __forceinline__ IVT<T> semicircle_angle_inner(IVT<T> la, IVT<T> r) {
  if constexpr (std::is_same<T, double>::value) {
    double ql, qu, yl, yu;
    div_pair2<false, true>(r.lb, add_ru(la.ub, r.lb), r.ub, add_rd(la.lb, r.ub), ql, qu);
    dasin_pair<false, true>(ql, qu, yl, yu);
    return IV(yl, yu);
  } else {
    return IVT<T>(dasin_rd(div_rd(r.lb, add_ru(la.ub, r.lb))), dasin_ru(div_ru(r.ub, add_rd(la.lb, r.ub))));
  }
}

// semicircle_angle_outer(r1), or `row` when it holds that angle for an r1 with
// the same bits (the reference-grid kernel computes it once per (la, r1) row;
// tightening leaves r1 unchanged for nearly every cell of the row).
template <class T>
__device__ __forceinline__ IVT<T> semicircle_outer_r1(IVT<T> r1, const IVT<T>* row) {
  if (row) return *row;
  return semicircle_angle_outer(r1);
}

// lower_bound_radius_single_placement.cu:26-61
__device__ __attribute__((noinline)) double lower_bound_radius_single_placement(double lb_lambda) {
  const double ub_rhs_factor = div_ru(critical_ratio, lb_pi);
  const double ub_rhs = mul_ru(ub_rhs_factor, add_ru(1.0, -mul_rd(lb_lambda, lb_lambda)));
  double ub = 0.25;
  double lb = 5.0e-324;
  for (;;) {
    double mid = 0.5 * (lb + ub);
    if (mid <= lb || mid >= ub) return ub;
    double r_sq = mul_rd(mid, mid);
    double ub_quot = div_ru(mid, add_rd(1.0, -mid));
    double ub_asin = dasin_ru(ub_quot);
    double lb_lhs = div_rd(r_sq, ub_asin);
    if (lb_lhs >= ub_rhs) ub = mid;
    else lb = mid;
  }
}

// ---- admissible ranges (a11), FP64 only ------------------------------------
// start_configuration.cu:58-66 = middle_configuration.cu:66-74 = end_configuration.cu:55-63
__device__ __forceinline__ IV compute_r1_range(IV la, double lbrsp) {
  double ub_by_lane_width = (0.5 * (1.0 - la)).ub;
  double lb_by_lane_width = (0.25 * (1.0 - la)).lb;
  return IV(lb_by_lane_width, ub_by_lane_width < lbrsp ? ub_by_lane_width : lbrsp);
}
__device__ __forceinline__ double quarter_lane_lb(IV la) { return (0.25 * (1.0 - la)).lb; }
// end_configuration.cu:65-69 and :148-152
__device__ __forceinline__ double end_lane_lb(IV la, double r_ub) {
  double l = mul_rd(0.5, add_rd(1.0, -add_ru(la.ub, mul_ru(2.0, r_ub))));
  return l < 0.0 ? 0.0 : l;
}
template <int F>
__device__ __forceinline__ IV r2_range(IV la, IV r1) {
  if constexpr (F == F_START_END_2D) return IV((0.5 * (1.0 - la - 2.0 * r1)).lb, r1.ub);  // start_configuration.cu:74-77
  else if constexpr (F == F_END_2D || F == F_END_3D) return IV(end_lane_lb(la, r1.ub), r1.ub);
  else return IV(quarter_lane_lb(la), r1.ub);  // start :68-71, middle :76-79
}
template <int F>
__device__ __forceinline__ IV r3_range(IV la, IV r2) {
  if constexpr (F == F_START_END_3D) return IV((0.5 * (1.0 - la - 2.0 * r2)).lb, r2.ub);  // start :84-87
  else if constexpr (F == F_END_3D) return IV(end_lane_lb(la, r2.ub), r2.ub);
  else return IV(quarter_lane_lb(la), r2.ub);  // start :79-82, middle :81-84
}

// ---- criteria ------------------------------------------------------------
template <class T>
__device__ __forceinline__ T ub_ratio_of(IVT<T> la) { return add_ru(T(1), -mul_rd(la.lb, la.lb)); }
template <class T>
__device__ __forceinline__ T lb_disk(IVT<T> r) { return mul_rd(LaneConst<T>::lbpi, mul_rd(r.lb, r.lb)); }
template <class T>
__device__ __forceinline__ T crit_of(T x) { return mul_ru(LaneConst<T>::crit, x); }

// start_configuration.cu:93-101
template <class T>
__device__ __forceinline__ bool handle_case_1(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> a12, const IVT<T>* a1_row) {
  T ub_ratio = ub_ratio_of(la);
  T lb_area1 = lb_disk(r1), lb_area2 = lb_disk(r2);
  IVT<T> a = semicircle_outer_r1(r1, a1_row) + a12;
  T ub_area_used = mul_ru(ub_ratio, mul_ru(T(0.5), a.ub));
  T lb_disk_area = add_rd(lb_area1, mul_rd(T(0.5), lb_area2));
  return lb_disk_area >= crit_of(ub_area_used);
}
// start_configuration.cu:103-112
template <class T>
__device__ __forceinline__ bool handle_case_2(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, IVT<T> a13,
                                              const IVT<T>* a1_row) {
  T ub_ratio = ub_ratio_of(la);
  T lb_area1 = lb_disk(r1), lb_area2 = lb_disk(r2), lb_area3 = lb_disk(r3);
  IVT<T> a = semicircle_outer_r1(r1, a1_row) + a13;
  T ub_area_used = mul_ru(ub_ratio, mul_ru(T(0.5), a.ub));
  T lb_disk_area = add_rd(add_rd(lb_area1, lb_area2), mul_rd(T(0.5), lb_area3));
  return lb_disk_area >= crit_of(ub_area_used);
}
// start_configuration.cu:169-188
template <class T>
__device__ __forceinline__ bool handle_2d_case(IVT<T> la, IVT<T> r1, IVT<T> r2, int& trap, const IVT<T>* a1_row) {
  IVT<T> a1 = semicircle_outer_r1(r1, a1_row);
  IVT<T> a2 = semicircle_angle_inner(la, r2);
  IVT<T> a12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  IVT<T> total_angle = a1 + a2 + a12;
  IVT<T> reuse_angle = a12 + a2 - a1;
  reuse_angle.tighten_lb(T(0));
  T ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(T(0.5), total_angle.ub));
  IVT<T> rho = T(1) - (la + T(2) * r2).square();
  rho.tighten_lb(T(0));
  T lb_area_reusable = mul_rd(mul_rd(T(0.5), reuse_angle.lb), rho.lb);
  T lb_disk_area = add_rd(lb_disk(r2), lb_disk(r1));
  return lb_disk_area >= crit_of(add_ru(ub_area_used, -lb_area_reusable));
}
// start_configuration.cu:227-265 (outer_outer(r1, r2) at :232 kept as in the reference)
template <class T>
__device__ __forceinline__ bool handle_3d_case(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, int& trap,
                                               const IVT<T>* a1_row) {
  using I = IVT<T>;
  I a1 = semicircle_outer_r1(r1, a1_row);
  I a2 = semicircle_angle_inner(la, r2);
  I a3 = semicircle_angle_outer(r3);
  I a12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  I a13 = collision_angle_outer_outer(r1, r2, trap);
  if (trap) return false;
  I a23 = collision_angle_inner_outer(la, r2, r3, trap);
  if (trap) return false;
  I b13_123 = a12 + a23;
  I b13(a13.lb < b13_123.lb ? b13_123.lb : a13.lb, a13.ub < b13_123.ub ? b13_123.ub : a13.ub);
  I a123 = a1 + b13 + a3;
  I a122 = a1 + a12 + a2;
  I a(a123.lb < a122.lb ? a122.lb : a123.lb, a123.ub < a122.ub ? a122.ub : a123.ub);
  T lb_reusable_angle = (a - a122).lb;
  if (lb_reusable_angle < T(0)) lb_reusable_angle = T(0);
  T lb_outer_rad = add_rd(T(1), -mul_ru(T(2), r3.ub));
  T lb_outer_rad_sq = mul_rd(lb_outer_rad, lb_outer_rad);
  T ub_inner_rad_sq = mul_ru(la.ub, la.ub);
  T lb_area_reused = mul_rd(mul_rd(T(0.5), lb_reusable_angle), add_rd(lb_outer_rad_sq, -ub_inner_rad_sq));
  if (lb_area_reused < T(0)) lb_area_reused = T(0);
  T ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(T(0.5), a.ub));
  ub_area_used = add_ru(ub_area_used, -lb_area_reused);
  T lb_disk_area = add_rd(lb_disk(r1), add_rd(lb_disk(r2), lb_disk(r3)));
  return lb_disk_area >= crit_of(ub_area_used);
}
// middle_configuration.cu:86-154 (the two cases share this tail)
template <class T>
__device__ __forceinline__ bool middle_area_checks(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, IVT<T> c12, IVT<T> c13,
                                                   IVT<T> c23) {
  T ub_ratio = ub_ratio_of(la);
  UB case_a = (c12 + c23 >= c13);
  T lb_area1 = lb_disk(r1), lb_area2 = lb_disk(r2);
  if (possibly(case_a)) {
    T ub_area_used = mul_ru(ub_ratio, mul_ru(T(0.5), c12.ub));
    T lb_disk_area = mul_rd(T(0.5), add_rd(lb_area1, lb_area2));
    if (lb_disk_area < crit_of(ub_area_used)) return false;
  }
  if (possibly(!case_a)) {
    T ub_area_used = mul_ru(ub_ratio, mul_ru(T(0.5), c13.ub));
    T lb_area3 = lb_disk(r3);
    T lb_disk_area = add_rd(mul_rd(T(0.5), add_rd(lb_area1, lb_area3)), lb_area2);
    if (lb_disk_area < crit_of(ub_area_used)) return false;
  }
  return true;
}
template <class T>
__device__ __forceinline__ bool handle_case_1_1(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, int& trap) {
  IVT<T> c12 = collision_angle_inner_outer(la, r1, r2, trap);
  if (trap) return false;
  IVT<T> c13 = collision_angle_inner_inner(la, r1, r3, trap);
  if (trap) return false;
  IVT<T> c23 = collision_angle_outer_inner(la, r2, r3, trap);
  if (trap) return false;
  return middle_area_checks(la, r1, r2, r3, c12, c13, c23);
}
template <class T>
__device__ __forceinline__ bool handle_case_2_1(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, int& trap) {
  IVT<T> c12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  IVT<T> c13 = collision_angle_outer_outer(r1, r3, trap);
  if (trap) return false;
  IVT<T> c23 = collision_angle_inner_outer(la, r2, r3, trap);
  if (trap) return false;
  return middle_area_checks(la, r1, r2, r3, c12, c13, c23);
}
// end_configuration.cu:71-85
template <class T>
__device__ __forceinline__ bool handle_inner_outer_2d(IVT<T> la, IVT<T> r1, IVT<T> r2, int& trap) {
  IVT<T> a = collision_angle_inner_outer(la, r1, r2, trap) + semicircle_angle_outer(r2);
  if (trap) return false;
  T ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(T(0.5), a.ub));
  IVT<T> rho = (T(1) - T(2) * r2).square() - la.square();
  rho.tighten_lb(T(0));
  T lb_alpha_reusable = (a - semicircle_angle_inner(la, r1)).lb;
  T lb_area_reusable = mul_rd(mul_rd(T(0.5), lb_alpha_reusable), rho.lb);
  T lb_disk_area = add_rd(lb_disk(r2), mul_rd(T(0.5), lb_disk(r1)));
  return lb_disk_area >= crit_of(add_ru(ub_area_used, -lb_area_reusable));
}
// end_configuration.cu:87-101
template <class T>
__device__ __forceinline__ bool handle_outer_inner_2d(IVT<T> la, IVT<T> r1, IVT<T> r2, int& trap,
                                                      const IVT<T>* a1_row) {
  IVT<T> a = collision_angle_outer_inner(la, r1, r2, trap) + semicircle_angle_inner(la, r2);
  if (trap) return false;
  T ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(T(0.5), a.ub));
  IVT<T> rho = T(1) - (la + T(2) * r2).square();
  rho.tighten_lb(T(0));
  T lb_alpha_reusable = (a - semicircle_outer_r1(r1, a1_row)).lb;
  T lb_area_reusable = mul_rd(mul_rd(T(0.5), lb_alpha_reusable), rho.lb);
  T lb_disk_area = add_rd(lb_disk(r2), mul_rd(T(0.5), lb_disk(r1)));
  return lb_disk_area >= crit_of(add_ru(ub_area_used, -lb_area_reusable));
}
// end_configuration.cu:154-181
template <class T>
__device__ __forceinline__ bool handle_r1_inner_3d(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, int& trap) {
  using I = IVT<T>;
  I a12 = collision_angle_inner_outer(la, r1, r2, trap);
  if (trap) return false;
  I a23 = collision_angle_outer_inner(la, r2, r3, trap);
  if (trap) return false;
  I a13 = collision_angle_inner_inner(la, r1, r3, trap);
  if (trap) return false;
  if (definitely(a12 + a23 >= a13)) return true;
  I a3 = semicircle_angle_inner(la, r3);
  I a = a13 + a3;
  I a2 = semicircle_angle_outer(r2);
  I reusable_alpha = a - (a12 + a2);
  I rho = T(1) - (la + T(2) * r3).square();
  rho.tighten_lb(T(0));
  T lb_area_reusable = mul_rd(mul_rd(T(0.5), reusable_alpha.lb), rho.lb);
  T ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(T(0.5), a.ub));
  T lb_disk_area = add_rd(add_rd(lb_disk(r2), lb_disk(r3)), mul_rd(T(0.5), lb_disk(r1)));
  return lb_disk_area >= crit_of(add_ru(ub_area_used, -lb_area_reusable));
}
// end_configuration.cu:183-210
template <class T>
__device__ __forceinline__ bool handle_r1_outer_3d(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, int& trap) {
  using I = IVT<T>;
  I a12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  I a23 = collision_angle_inner_outer(la, r2, r3, trap);
  if (trap) return false;
  I a13 = collision_angle_outer_outer(r1, r3, trap);
  if (trap) return false;
  if (definitely(a12 + a23 >= a13)) return true;
  I a3 = semicircle_angle_outer(r3);
  I a = a13 + a3;
  I a2 = semicircle_angle_inner(la, r2);
  I reusable_alpha = a - (a12 + a2);
  I rho = (T(1) - T(2) * r3).square() - la.square();
  rho.tighten_lb(T(0));
  T lb_area_reusable = mul_rd(mul_rd(T(0.5), reusable_alpha.lb), rho.lb);
  T ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(T(0.5), a.ub));
  T lb_disk_area = add_rd(add_rd(lb_disk(r2), lb_disk(r3)), mul_rd(T(0.5), lb_disk(r1)));
  return lb_disk_area >= crit_of(add_ru(ub_area_used, -lb_area_reusable));
}

// The gated criteria of family F on tightened cells.  Returns 0 (all hold),
// the 1-based failing criterion, or 16 + trap code.  Inline: out of line, the
// call would save and restore ~40 callee-saved VGPRs per lane through scratch
// (measured: HBM traffic of the round kernel x1.8 for no gain in time).
// a1_row: semicircle_angle_outer(r1) when the caller has it (null otherwise).
// Two-check families (middle, end 2-disk, end 3-disk): the checks in the
// reference's order, or with second_first (wave-uniform; the learned order)
// the second check first.  The result names the first failing check in the
// order evaluated, so it differs from the reference's only where both fail.
template <class C1, class C2>
__device__ __forceinline__ int two_checks(bool second_first, C1 check1, C2 check2) {
  int trap = 0;
#pragma unroll 1
  for (int k = 0; k < 2; ++k) {
    const bool second = (k == 0) == second_first;
    const bool ok = second ? check2(trap) : check1(trap);
    if (trap) return 16 + trap;
    if (!ok) return second ? 2 : 1;
  }
  return 0;
}

template <int F, class T>
__device__ __forceinline__ int criteria(IVT<T> la, IVT<T> r1, IVT<T> r2, IVT<T> r3, const IVT<T>* a1_row = nullptr,
                                        bool second_first = false) {
  int trap = 0;
  if constexpr (F == F_START) {  // start_configuration.cu:142-162
    IVT<T> a12 = collision_angle_outer_inner(la, r1, r2, trap);
    if (trap) return 16 + trap;
    IVT<T> a13 = collision_angle_outer_outer(r1, r3, trap);
    if (trap) return 16 + trap;
    IVT<T> a23 = collision_angle_inner_outer(la, r2, r3, trap);
    if (trap) return 16 + trap;
    UB case1 = (a12 + a23 > a13);
    if (possibly(case1) && !handle_case_1(la, r1, r2, a12, a1_row)) return 1;
    if (possibly(!case1) && !handle_case_2(la, r1, r2, r3, a13, a1_row)) return 2;
    return 0;
  } else if constexpr (F == F_START_END_2D) {
    bool ok = handle_2d_case(la, r1, r2, trap, a1_row);
    if (trap) return 16 + trap;
    return ok ? 0 : 1;
  } else if constexpr (F == F_START_END_3D) {
    bool ok = handle_3d_case(la, r1, r2, r3, trap, a1_row);
    if (trap) return 16 + trap;
    return ok ? 0 : 1;
  } else if constexpr (F == F_MIDDLE) {
    if (second_first)
      return two_checks(
          true, [&](int& tr) { return handle_case_1_1(la, r1, r2, r3, tr); },
          [&](int& tr) { return handle_case_2_1(la, r1, r2, r3, tr); });
    bool ok = handle_case_1_1(la, r1, r2, r3, trap);
    if (trap) return 16 + trap;
    if (!ok) return 1;
    ok = handle_case_2_1(la, r1, r2, r3, trap);
    if (trap) return 16 + trap;
    return ok ? 0 : 2;
  } else if constexpr (F == F_END_2D) {
    if (second_first)
      return two_checks(
          true, [&](int& tr) { return handle_inner_outer_2d(la, r1, r2, tr); },
          [&](int& tr) { return handle_outer_inner_2d(la, r1, r2, tr, a1_row); });
    bool ok = handle_inner_outer_2d(la, r1, r2, trap);
    if (trap) return 16 + trap;
    if (!ok) return 1;
    ok = handle_outer_inner_2d(la, r1, r2, trap, a1_row);
    if (trap) return 16 + trap;
    return ok ? 0 : 2;
  } else {
    if (second_first)
      return two_checks(
          true, [&](int& tr) { return handle_r1_inner_3d(la, r1, r2, r3, tr); },
          [&](int& tr) { return handle_r1_outer_3d(la, r1, r2, r3, tr); });
    bool ok = handle_r1_inner_3d(la, r1, r2, r3, trap);
    if (trap) return 16 + trap;
    if (!ok) return 1;
    ok = handle_r1_outer_3d(la, r1, r2, r3, trap);
    if (trap) return 16 + trap;
    return ok ? 0 : 2;
  }
}

// ---- kernel body (tighten, empty, fit check, gated criteria) ----------------
// start_configuration.cu:127-162, :200-212, :280-296; middle_configuration.cu:168-202;
// end_configuration.cu:110-143, :226-265.
// PREFILTER (and `pre` at run time): the criteria are first evaluated on the
// FP32 outward enclosure of the tightened cell; a cell FP32 proves is PROVEN,
// every other cell is decided by the FP64 evaluation.  `*fp32_hit` is set when
// FP32 settled the cell, `*fp32_tried` when the FP32 criteria ran.
// a1_row (reference-grid kernel): semicircle_angle_outer(r1_), used where
// tightening leaves r1 = r1_ bit for bit.
template <int F, bool PREFILTER = false>
__device__ __forceinline__ uint8_t eval_kernel_body(IV la, IV r1_, IV r2_, IV r3, bool pre = true,
                                                    bool* fp32_hit = nullptr, bool* fp32_tried = nullptr,
                                                    const IV* a1_row = nullptr, bool second_first = false) {
  IV r2 = r2_;
  if constexpr (is_3d(F)) r2.tighten_lb(r3.lb);
  IV r1 = r1_;
  r1.tighten_lb(r2.lb);
  if (r1.empty() || r2.empty()) return vbyte(V_EMPTY, 0, 0);
  int nofit = 0;
  if constexpr (F == F_START || F == F_MIDDLE || F == F_END_2D || F == F_END_3D) {
    if (mul_ru(2.0, r1.ub) >= add_rd(1.0, -la.ub)) {
      if constexpr (F != F_START) return vbyte(V_NOFIT, 0, 0);
      nofit = 1;
    }
  }
  if constexpr (PREFILTER) {
    if (pre) {
      if (fp32_tried) *fp32_tried = true;
      int c32 = criteria<F, float>(to_f32(la), to_f32(r1), to_f32(r2), to_f32(r3));
      if (c32 == 0) {
        if (fp32_hit) *fp32_hit = true;
        return vbyte(V_PROVEN, nofit, 0);
      }
    }
  }
  const IV* a1 = (a1_row && dbits(r1.lb) == dbits(r1_.lb) && dbits(r1.ub) == dbits(r1_.ub)) ? a1_row : nullptr;
  int c = criteria<F, double>(la, r1, r2, r3, a1, second_first);
  if (c >= 16) return vbyte(V_TRAP, nofit, c - 16);
  if (c) return vbyte(V_CRITICAL, nofit, c);
  return vbyte(V_PROVEN, nofit, 0);
}

// The reference's output_criticals mode on one cell: the kernels keep
// evaluating after a failing criterion and print one line per failure
// (start_configuration.cu:142-161, :202-210, :284-292; middle_configuration.cu:
// 180-200; end_configuration.cu:122-142, :243-263).  Returns the mask of
// failing criteria (bit k-1: criterion k), or 16 + trap code when a check of
// any criterion traps; *nofit_flag: the start family's "possibly does not fit"
// line (start_configuration.cu:131-135, printed before the criteria run);
// *early: V_EMPTY, or V_NOFIT for the families that trap at the fit check,
// when the cell is decided before the criteria (mask 0), else -1.
template <int F>
__device__ __forceinline__ int eval_kernel_body_all(IV la, IV r1_, IV r2_, IV r3, int* nofit_flag, int* early) {
  IV r2 = r2_;
  if constexpr (is_3d(F)) r2.tighten_lb(r3.lb);
  IV r1 = r1_;
  r1.tighten_lb(r2.lb);
  *nofit_flag = 0;
  *early = -1;
  if (r1.empty() || r2.empty()) {
    *early = V_EMPTY;
    return 0;
  }
  if constexpr (F == F_START || F == F_MIDDLE || F == F_END_2D || F == F_END_3D) {
    if (mul_ru(2.0, r1.ub) >= add_rd(1.0, -la.ub)) {
      if constexpr (F != F_START) {
        *early = V_NOFIT;
        return 0;
      }
      *nofit_flag = 1;
    }
  }
  int trap = 0, mask = 0;
  if constexpr (F == F_START) {
    IV a12 = collision_angle_outer_inner(la, r1, r2, trap);
    if (trap) return 16 + trap;
    IV a13 = collision_angle_outer_outer(r1, r3, trap);
    if (trap) return 16 + trap;
    IV a23 = collision_angle_inner_outer(la, r2, r3, trap);
    if (trap) return 16 + trap;
    UB case1 = (a12 + a23 > a13);
    if (possibly(case1) && !handle_case_1<double>(la, r1, r2, a12, nullptr)) mask |= 1;
    if (possibly(!case1) && !handle_case_2<double>(la, r1, r2, r3, a13, nullptr)) mask |= 2;
  } else if constexpr (F == F_START_END_2D) {
    const bool ok = handle_2d_case<double>(la, r1, r2, trap, nullptr);
    if (trap) return 16 + trap;
    mask = ok ? 0 : 1;
  } else if constexpr (F == F_START_END_3D) {
    const bool ok = handle_3d_case<double>(la, r1, r2, r3, trap, nullptr);
    if (trap) return 16 + trap;
    mask = ok ? 0 : 1;
  } else {
    bool ok1, ok2;
    if constexpr (F == F_MIDDLE) {
      ok1 = handle_case_1_1<double>(la, r1, r2, r3, trap);
      if (trap) return 16 + trap;
      ok2 = handle_case_2_1<double>(la, r1, r2, r3, trap);
    } else if constexpr (F == F_END_2D) {
      ok1 = handle_inner_outer_2d<double>(la, r1, r2, trap);
      if (trap) return 16 + trap;
      ok2 = handle_outer_inner_2d<double>(la, r1, r2, trap, nullptr);
    } else {
      ok1 = handle_r1_inner_3d<double>(la, r1, r2, r3, trap);
      if (trap) return 16 + trap;
      ok2 = handle_r1_outer_3d<double>(la, r1, r2, r3, trap);
    }
    if (trap) return 16 + trap;
    mask = (ok1 ? 0 : 1) | (ok2 ? 0 : 2);
  }
  return mask;
}

// r3 from a position interval p in [0,1] over the admissible r3 range; for a
// root cell p = [k/n, (k+1)/n] this reproduces get_subinterval(range, k, n).
__device__ __forceinline__ IV r3_from_position(IV range, IV p) {
  double w = range.ub - range.lb;
  double l = __builtin_fma(w, p.lb, range.lb);
  double u = (p.ub == 1.0) ? range.ub : __builtin_fma(w, p.ub, range.lb);
  return IV(l, u);
}

// Synthetic cell (la, r1, r2, p3) of family F; lbrsp = lower_bound_radius_single_placement(la.lb).
// second_first: the learned check order (two-check families, wave-uniform).
template <int F, bool PREFILTER = false>
__device__ __forceinline__ uint8_t eval_synthetic(IV la, IV r1, IV r2, IV p3, bool clip, double lbrsp,
                                                  bool pre = true, bool* fp32_hit = nullptr,
                                                  bool* fp32_tried = nullptr, bool second_first = false) {
  IV r1c = r1, r2c = r2;
  if (clip) {
    r1c = r1.intersect(compute_r1_range(la, lbrsp));
    if (r1c.empty()) return vbyte(V_OOD, 0, 0);
    r2c = r2.intersect(r2_range<F>(la, r1c));
    if (r2c.empty()) return vbyte(V_OOD, 0, 0);
  }
  IV r3(0.0, 0.0);
  if constexpr (is_3d(F)) {
    IV range = r3_range<F>(la, r2c);
    if (range.empty()) return vbyte(V_OOD, 0, 0);
    r3 = r3_from_position(range, p3);
  }
  return eval_kernel_body<F, PREFILTER>(la, r1c, r2c, r3, pre, fp32_hit, fp32_tried, nullptr, second_first);
}

}  // namespace cpl
