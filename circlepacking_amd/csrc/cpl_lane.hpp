// Lane geometry, criteria and the per-cell kernel body on the device.
//
// Follows lane_packing/collision_angles.cu, lower_bound_radius_single_placement.cu
// and the three *_configuration.cu files of the reference formula for formula
// (each function cites its source lines).  Differences in mechanism only:
//  * trap() becomes a trap code in `trap` (first one wins; callers stop at once);
//  * the kernel body returns a verdict byte instead of calling store_min/printf.
#pragma once

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

// Interval asin (interval.cuh:617-624).
__device__ __forceinline__ IV asin_iv(IV x, int& trap) {
  if (x.lb < -1.0 || x.ub > 1.0) {
    if (!trap) trap = T_ASIN;
    return IV(0.0, 0.0);
  }
  return IV(dasin_rd(x.lb), dasin_ru(x.ub));
}

// collision_angles.cu:29-31
__device__ __forceinline__ UB is_obtuse(IV hyp, IV s1, IV s2) { return hyp.square() > s1.square() + s2.square(); }

// collision_angles.cu:37-117
__device__ __attribute__((noinline)) IV compute_collision_angle(IV lc, IV sc, IV ab, int& trap) {
  lc.tighten_lb(sc.lb);
  sc.tighten_ub(lc.ub);
  IV result(0.0, -1.0);
  UB obt = is_obtuse(lc, sc, ab);
  IV lc_sq = lc.square();
  IV sc_sq = sc.square();
  IV add_term = (lc_sq - sc_sq) / ab;
  add_term.tighten_lb(0.0);

  if (possibly(obt)) {
    IV lx = 0.5 * (add_term + ab);
    IV sx = 0.5 * (add_term - ab);
    sx.tighten_lb(0.0);
    sx.tighten_ub(sc.ub);
    lx.tighten_ub(lc.ub);
    if (sx.empty() || lx.empty()) {
      if (definitely(obt)) { trap = T_OBT_X; return result; }
    } else {
      IV lquot = lx / lc;
      lquot.tighten_ub(1.0);
      IV squot = sx / sc;
      squot.tighten_ub(1.0);
      if (lquot.empty() || squot.empty()) {
        if (definitely(obt)) { trap = T_OBT_Q; return result; }
      } else {
        IV alpha = asin_iv(lquot, trap);
        if (trap) return result;
        IV beta = asin_iv(squot, trap);
        if (trap) return result;
        result = alpha - beta;
        result.tighten_lb(0.0);
      }
    }
  }

  if (!definitely(obt)) {
    IV lx = 0.5 * (ab + add_term);
    IV sx = 0.5 * (ab - add_term);
    sx.tighten_lb(0.0);
    if (sx.empty()) {
      if (!possibly(obt)) { trap = T_ACU_X; return result; }
    } else {
      IV alpha = asin_iv(lx / lc, trap);
      if (trap) return result;
      IV beta = asin_iv(sx / sc, trap);
      if (trap) return result;
      beta.tighten_lb(0.0);
      IV r = alpha + beta;
      if (result.empty()) result = r;
      else result.do_join(r);
    }
  }
  return result;
}

// collision_angles.cu:119-142
__device__ __attribute__((noinline)) IV collision_angle_inner_inner(IV la, IV ra, IV rb, int& trap) {
  double lb, ub;
  {
    IV lc = la.lb + IV(ra.ub, ra.ub);
    IV sc = la.lb + IV(rb.ub, rb.ub);
    IV ab = IV(ra.ub, ra.ub) + rb.ub;
    ub = compute_collision_angle(lc, sc, ab, trap).ub;
    if (trap) return IV(0, 0);
  }
  {
    IV lc = la.ub + IV(ra.lb, ra.lb);
    IV sc = la.ub + IV(rb.lb, rb.lb);
    IV ab = IV(ra.lb, ra.lb) + rb.lb;
    if (ab.lb <= 0.0) lb = 0.0;
    else lb = compute_collision_angle(lc, sc, ab, trap).lb;
  }
  return IV(lb, ub);
}

// collision_angles.cu:144-159 (outer_inner: ca = 1 - ra, cb = la + rb) and
// :176-191 (inner_outer: ca = la + ra, cb = 1 - rb).
__device__ __forceinline__ IV do_collision_angle(IV ca, IV cb, IV ab, int& trap) {
  UB ca_longer = (cb <= ca);
  if (definitely(ca_longer)) return compute_collision_angle(ca, cb, ab, trap);
  if (definitely(!ca_longer)) return compute_collision_angle(cb, ca, ab, trap);
  IV result = compute_collision_angle(ca, cb, ab, trap);
  if (trap) return result;
  result.do_join(compute_collision_angle(cb, ca, ab, trap));
  return result;
}

// collision_angles.cu:161-174
__device__ __attribute__((noinline)) IV collision_angle_outer_inner(IV la, IV ra, IV rb, int& trap) {
  double lb, ub;
  if (possibly(2.0 * (ra + rb) <= 1.0 - la)) {
    lb = 0.0;
  } else {
    IV pa(ra.lb, ra.lb), pb(rb.lb, rb.lb);
    lb = do_collision_angle(1.0 - pa, la + pb, pa + pb, trap).lb;
    if (trap) return IV(0, 0);
  }
  IV pa(ra.ub, ra.ub), pb(rb.ub, rb.ub);
  ub = do_collision_angle(1.0 - pa, la + pb, pa + pb, trap).ub;
  return IV(lb, ub);
}

// collision_angles.cu:193-205
__device__ __attribute__((noinline)) IV collision_angle_inner_outer(IV la, IV ra, IV rb, int& trap) {
  double lb, ub;
  if (possibly(2.0 * (ra + rb) <= 1.0 - la)) {
    lb = 0.0;
  } else {
    IV pa(ra.lb, ra.lb), pb(rb.lb, rb.lb);
    lb = do_collision_angle(la + pa, 1.0 - pb, pa + pb, trap).lb;
    if (trap) return IV(0, 0);
  }
  IV pa(ra.ub, ra.ub), pb(rb.ub, rb.ub);
  ub = do_collision_angle(la + pa, 1.0 - pb, pa + pb, trap).ub;
  return IV(lb, ub);
}

// collision_angles.cu:207-231
__device__ __attribute__((noinline)) IV collision_angle_outer_outer(IV ra, IV rb, int& trap) {
  double lb, ub;
  {
    IV lc = IV(1.0, 1.0) - rb.ub;
    IV sc = IV(1.0, 1.0) - ra.ub;
    IV ab = IV(ra.ub, ra.ub) + rb.ub;
    ub = compute_collision_angle(lc, sc, ab, trap).ub;
    if (trap) return IV(0, 0);
  }
  {
    IV lc = IV(1.0, 1.0) - rb.lb;
    IV sc = IV(1.0, 1.0) - ra.lb;
    IV ab = IV(ra.lb, ra.lb) + rb.lb;
    if (ab.lb <= 0.0) lb = 0.0;
    else lb = compute_collision_angle(lc, sc, ab, trap).lb;
  }
  return IV(lb, ub);
}

// collision_angles.cu:233-238
__device__ __forceinline__ IV semicircle_angle_outer(IV r) {
  return IV(dasin_rd(div_rd(r.lb, add_ru(1.0, -r.lb))), dasin_ru(div_ru(r.ub, add_rd(1.0, -r.ub))));
}
// collision_angles.cu:240-246
__device__ __forceinline__ IV semicircle_angle_inner(IV la, IV r) {
  return IV(dasin_rd(div_rd(r.lb, add_ru(la.ub, r.lb))), dasin_ru(div_ru(r.ub, add_rd(la.lb, r.ub))));
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

// ---- admissible ranges (a11) ------------------------------------------------
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
__device__ __forceinline__ double ub_ratio_of(IV la) { return add_ru(1.0, -mul_rd(la.lb, la.lb)); }
__device__ __forceinline__ double lb_disk(IV r) { return mul_rd(lb_pi, mul_rd(r.lb, r.lb)); }

// start_configuration.cu:93-101
__device__ __forceinline__ bool handle_case_1(IV la, IV r1, IV r2, IV a12) {
  double ub_ratio = ub_ratio_of(la);
  double lb_area1 = lb_disk(r1), lb_area2 = lb_disk(r2);
  IV a = semicircle_angle_outer(r1) + a12;
  double ub_area_used = mul_ru(ub_ratio, mul_ru(0.5, a.ub));
  double lb_disk_area = add_rd(lb_area1, mul_rd(0.5, lb_area2));
  return lb_disk_area >= mul_ru(critical_ratio, ub_area_used);
}
// start_configuration.cu:103-112
__device__ __forceinline__ bool handle_case_2(IV la, IV r1, IV r2, IV r3, IV a13) {
  double ub_ratio = ub_ratio_of(la);
  double lb_area1 = lb_disk(r1), lb_area2 = lb_disk(r2), lb_area3 = lb_disk(r3);
  IV a = semicircle_angle_outer(r1) + a13;
  double ub_area_used = mul_ru(ub_ratio, mul_ru(0.5, a.ub));
  double lb_disk_area = add_rd(add_rd(lb_area1, lb_area2), mul_rd(0.5, lb_area3));
  return lb_disk_area >= mul_ru(critical_ratio, ub_area_used);
}
// start_configuration.cu:169-188
__device__ __forceinline__ bool handle_2d_case(IV la, IV r1, IV r2, int& trap) {
  IV a1 = semicircle_angle_outer(r1);
  IV a2 = semicircle_angle_inner(la, r2);
  IV a12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  IV total_angle = a1 + a2 + a12;
  IV reuse_angle = a12 + a2 - a1;
  reuse_angle.tighten_lb(0.0);
  double ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(0.5, total_angle.ub));
  IV rho = 1.0 - (la + 2.0 * r2).square();
  rho.tighten_lb(0.0);
  double lb_area_reusable = mul_rd(mul_rd(0.5, reuse_angle.lb), rho.lb);
  double lb_disk_area = add_rd(lb_disk(r2), lb_disk(r1));
  return lb_disk_area >= mul_ru(critical_ratio, add_ru(ub_area_used, -lb_area_reusable));
}
// start_configuration.cu:227-265 (outer_outer(r1, r2) at :232 kept as in the reference)
__device__ __forceinline__ bool handle_3d_case(IV la, IV r1, IV r2, IV r3, int& trap) {
  IV a1 = semicircle_angle_outer(r1);
  IV a2 = semicircle_angle_inner(la, r2);
  IV a3 = semicircle_angle_outer(r3);
  IV a12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  IV a13 = collision_angle_outer_outer(r1, r2, trap);
  if (trap) return false;
  IV a23 = collision_angle_inner_outer(la, r2, r3, trap);
  if (trap) return false;
  IV b13_123 = a12 + a23;
  IV b13(a13.lb < b13_123.lb ? b13_123.lb : a13.lb, a13.ub < b13_123.ub ? b13_123.ub : a13.ub);
  IV a123 = a1 + b13 + a3;
  IV a122 = a1 + a12 + a2;
  IV a(a123.lb < a122.lb ? a122.lb : a123.lb, a123.ub < a122.ub ? a122.ub : a123.ub);
  double lb_reusable_angle = (a - a122).lb;
  if (lb_reusable_angle < 0.0) lb_reusable_angle = 0.0;
  double lb_outer_rad = add_rd(1.0, -mul_ru(2.0, r3.ub));
  double lb_outer_rad_sq = mul_rd(lb_outer_rad, lb_outer_rad);
  double ub_inner_rad_sq = mul_ru(la.ub, la.ub);
  double lb_area_reused = mul_rd(mul_rd(0.5, lb_reusable_angle), add_rd(lb_outer_rad_sq, -ub_inner_rad_sq));
  if (lb_area_reused < 0.0) lb_area_reused = 0.0;
  double ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(0.5, a.ub));
  ub_area_used = add_ru(ub_area_used, -lb_area_reused);
  double lb_disk_area = add_rd(lb_disk(r1), add_rd(lb_disk(r2), lb_disk(r3)));
  return lb_disk_area >= mul_ru(critical_ratio, ub_area_used);
}
// middle_configuration.cu:86-154 (the two cases share this tail)
__device__ __forceinline__ bool middle_area_checks(IV la, IV r1, IV r2, IV r3, IV c12, IV c13, IV c23) {
  double ub_ratio = ub_ratio_of(la);
  UB case_a = (c12 + c23 >= c13);
  double lb_area1 = lb_disk(r1), lb_area2 = lb_disk(r2);
  if (possibly(case_a)) {
    double ub_area_used = mul_ru(ub_ratio, mul_ru(0.5, c12.ub));
    double lb_disk_area = mul_rd(0.5, add_rd(lb_area1, lb_area2));
    if (lb_disk_area < mul_ru(critical_ratio, ub_area_used)) return false;
  }
  if (possibly(!case_a)) {
    double ub_area_used = mul_ru(ub_ratio, mul_ru(0.5, c13.ub));
    double lb_area3 = lb_disk(r3);
    double lb_disk_area = add_rd(mul_rd(0.5, add_rd(lb_area1, lb_area3)), lb_area2);
    if (lb_disk_area < mul_ru(critical_ratio, ub_area_used)) return false;
  }
  return true;
}
__device__ __forceinline__ bool handle_case_1_1(IV la, IV r1, IV r2, IV r3, int& trap) {
  IV c12 = collision_angle_inner_outer(la, r1, r2, trap);
  if (trap) return false;
  IV c13 = collision_angle_inner_inner(la, r1, r3, trap);
  if (trap) return false;
  IV c23 = collision_angle_outer_inner(la, r2, r3, trap);
  if (trap) return false;
  return middle_area_checks(la, r1, r2, r3, c12, c13, c23);
}
__device__ __forceinline__ bool handle_case_2_1(IV la, IV r1, IV r2, IV r3, int& trap) {
  IV c12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  IV c13 = collision_angle_outer_outer(r1, r3, trap);
  if (trap) return false;
  IV c23 = collision_angle_inner_outer(la, r2, r3, trap);
  if (trap) return false;
  return middle_area_checks(la, r1, r2, r3, c12, c13, c23);
}
// end_configuration.cu:71-85
__device__ __forceinline__ bool handle_inner_outer_2d(IV la, IV r1, IV r2, int& trap) {
  IV a = collision_angle_inner_outer(la, r1, r2, trap) + semicircle_angle_outer(r2);
  if (trap) return false;
  double ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(0.5, a.ub));
  IV rho = (1.0 - 2.0 * r2).square() - la.square();
  rho.tighten_lb(0.0);
  double lb_alpha_reusable = (a - semicircle_angle_inner(la, r1)).lb;
  double lb_area_reusable = mul_rd(mul_rd(0.5, lb_alpha_reusable), rho.lb);
  double lb_disk_area = add_rd(lb_disk(r2), mul_rd(0.5, lb_disk(r1)));
  return lb_disk_area >= mul_ru(critical_ratio, add_ru(ub_area_used, -lb_area_reusable));
}
// end_configuration.cu:87-101
__device__ __forceinline__ bool handle_outer_inner_2d(IV la, IV r1, IV r2, int& trap) {
  IV a = collision_angle_outer_inner(la, r1, r2, trap) + semicircle_angle_inner(la, r2);
  if (trap) return false;
  double ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(0.5, a.ub));
  IV rho = 1.0 - (la + 2.0 * r2).square();
  rho.tighten_lb(0.0);
  double lb_alpha_reusable = (a - semicircle_angle_outer(r1)).lb;
  double lb_area_reusable = mul_rd(mul_rd(0.5, lb_alpha_reusable), rho.lb);
  double lb_disk_area = add_rd(lb_disk(r2), mul_rd(0.5, lb_disk(r1)));
  return lb_disk_area >= mul_ru(critical_ratio, add_ru(ub_area_used, -lb_area_reusable));
}
// end_configuration.cu:154-181
__device__ __forceinline__ bool handle_r1_inner_3d(IV la, IV r1, IV r2, IV r3, int& trap) {
  IV a12 = collision_angle_inner_outer(la, r1, r2, trap);
  if (trap) return false;
  IV a23 = collision_angle_outer_inner(la, r2, r3, trap);
  if (trap) return false;
  IV a13 = collision_angle_inner_inner(la, r1, r3, trap);
  if (trap) return false;
  if (definitely(a12 + a23 >= a13)) return true;
  IV a3 = semicircle_angle_inner(la, r3);
  IV a = a13 + a3;
  IV a2 = semicircle_angle_outer(r2);
  IV reusable_alpha = a - (a12 + a2);
  IV rho = 1.0 - (la + 2.0 * r3).square();
  rho.tighten_lb(0.0);
  double lb_area_reusable = mul_rd(mul_rd(0.5, reusable_alpha.lb), rho.lb);
  double ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(0.5, a.ub));
  double lb_disk_area = add_rd(add_rd(lb_disk(r2), lb_disk(r3)), mul_rd(0.5, lb_disk(r1)));
  return lb_disk_area >= mul_ru(critical_ratio, add_ru(ub_area_used, -lb_area_reusable));
}
// end_configuration.cu:183-210
__device__ __forceinline__ bool handle_r1_outer_3d(IV la, IV r1, IV r2, IV r3, int& trap) {
  IV a12 = collision_angle_outer_inner(la, r1, r2, trap);
  if (trap) return false;
  IV a23 = collision_angle_inner_outer(la, r2, r3, trap);
  if (trap) return false;
  IV a13 = collision_angle_outer_outer(r1, r3, trap);
  if (trap) return false;
  if (definitely(a12 + a23 >= a13)) return true;
  IV a3 = semicircle_angle_outer(r3);
  IV a = a13 + a3;
  IV a2 = semicircle_angle_inner(la, r2);
  IV reusable_alpha = a - (a12 + a2);
  IV rho = (1.0 - 2.0 * r3).square() - la.square();
  rho.tighten_lb(0.0);
  double lb_area_reusable = mul_rd(mul_rd(0.5, reusable_alpha.lb), rho.lb);
  double ub_area_used = mul_ru(ub_ratio_of(la), mul_ru(0.5, a.ub));
  double lb_disk_area = add_rd(add_rd(lb_disk(r2), lb_disk(r3)), mul_rd(0.5, lb_disk(r1)));
  return lb_disk_area >= mul_ru(critical_ratio, add_ru(ub_area_used, -lb_area_reusable));
}

// ---- kernel body (tighten, empty, fit check, gated criteria) ----------------
// start_configuration.cu:127-162, :200-212, :280-296; middle_configuration.cu:168-202;
// end_configuration.cu:110-143, :226-265.
template <int F>
__device__ __forceinline__ uint8_t eval_kernel_body(IV la, IV r1_, IV r2_, IV r3) {
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
  int trap = 0;
  uint8_t v;
  if constexpr (F == F_START) {
    IV a12 = collision_angle_outer_inner(la, r1, r2, trap);
    if (trap) return vbyte(V_TRAP, nofit, trap);
    IV a13 = collision_angle_outer_outer(r1, r3, trap);
    if (trap) return vbyte(V_TRAP, nofit, trap);
    IV a23 = collision_angle_inner_outer(la, r2, r3, trap);
    if (trap) return vbyte(V_TRAP, nofit, trap);
    UB case1 = (a12 + a23 > a13);
    if (possibly(case1) && !handle_case_1(la, r1, r2, a12)) return vbyte(V_CRITICAL, nofit, 1);
    if (possibly(!case1) && !handle_case_2(la, r1, r2, r3, a13)) return vbyte(V_CRITICAL, nofit, 2);
    v = vbyte(V_PROVEN, nofit, 0);
  } else if constexpr (F == F_START_END_2D) {
    bool ok = handle_2d_case(la, r1, r2, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    v = ok ? vbyte(V_PROVEN, 0, 0) : vbyte(V_CRITICAL, 0, 1);
  } else if constexpr (F == F_START_END_3D) {
    bool ok = handle_3d_case(la, r1, r2, r3, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    v = ok ? vbyte(V_PROVEN, 0, 0) : vbyte(V_CRITICAL, 0, 1);
  } else if constexpr (F == F_MIDDLE) {
    bool ok = handle_case_1_1(la, r1, r2, r3, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    if (!ok) return vbyte(V_CRITICAL, 0, 1);
    ok = handle_case_2_1(la, r1, r2, r3, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    v = ok ? vbyte(V_PROVEN, 0, 0) : vbyte(V_CRITICAL, 0, 2);
  } else if constexpr (F == F_END_2D) {
    bool ok = handle_inner_outer_2d(la, r1, r2, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    if (!ok) return vbyte(V_CRITICAL, 0, 1);
    ok = handle_outer_inner_2d(la, r1, r2, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    v = ok ? vbyte(V_PROVEN, 0, 0) : vbyte(V_CRITICAL, 0, 2);
  } else {
    bool ok = handle_r1_inner_3d(la, r1, r2, r3, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    if (!ok) return vbyte(V_CRITICAL, 0, 1);
    ok = handle_r1_outer_3d(la, r1, r2, r3, trap);
    if (trap) return vbyte(V_TRAP, 0, trap);
    v = ok ? vbyte(V_PROVEN, 0, 0) : vbyte(V_CRITICAL, 0, 2);
  }
  return v;
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
template <int F>
__device__ __forceinline__ uint8_t eval_synthetic(IV la, IV r1, IV r2, IV p3, bool clip, double lbrsp) {
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
  return eval_kernel_body<F>(la, r1c, r2c, r3);
}

}  // namespace cpl
