// Interval sin, cos, acos, sqrt, reciprocal and scalar division: the rest of
// the reference's interval surface (interval.cuh:217-350, 357-438, 440-615).
// Not on the lane-packing path (no call site there); provided so that code
// written against algcuda::Interval<double> finds every operation.
//
// CUDA's ::sin, ::cos, ::acos are closed source and not reproduced: sin_rn,
// cos_rn and acos_rn below are our own round-to-nearest implementations
// (gen_trig_poly.py: three-part Cody-Waite reduction by pi/2, Chebyshev
// kernels; acos from the arcsine kernel of cpl_asin.hpp), and the directed
// versions step SC_STEPS ulps outward as the reference does (interval.cuh:
// 376-418), with SC_STEPS set from the measured error of our functions.
#pragma once

#include "cpl_asin.hpp"

namespace cpl {

constexpr int SC_STEPS = 3;    // outward ulp steps of dsin/dcos/dacos (measured error < 1 ulp)

namespace trig {
constexpr double lb_pi_two = 1.5707963267948966, ub_pi_two = 1.5707963267948968;  // interval.cuh:358-365
constexpr double lb_pi = 3.141592653589793, ub_pi = 3.1415926535897936;
constexpr double lb_3pi_two = 4.71238898038469, ub_3pi_two = 4.712388980384691;
constexpr double lb_two_pi = 6.283185307179586, ub_two_pi = 6.283185307179587;
constexpr double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54, P3 = -0x1.f1976b7ed8fbcp-110;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
}  // namespace trig

__device__ __forceinline__ double sin_poly(double z) {
  double p = 0x1.9242bc8544292p-49;
  p = __builtin_fma(p, z, -0x1.ae7c53ccfc679p-41);
  p = __builtin_fma(p, z, 0x1.61245fb507691p-33);
  p = __builtin_fma(p, z, -0x1.ae64567e73219p-26);
  p = __builtin_fma(p, z, 0x1.71de3a556c2c8p-19);
  p = __builtin_fma(p, z, -0x1.a01a01a01a019p-13);
  p = __builtin_fma(p, z, 0x1.1111111111111p-7);
  p = __builtin_fma(p, z, -0x1.5555555555555p-3);
  return p;
}
__device__ __forceinline__ double cos_poly(double z) {
  double p = -0x1.65d2c0a3f787dp-53;
  p = __builtin_fma(p, z, 0x1.ae7ce8e10b0f8p-45);
  p = __builtin_fma(p, z, -0x1.93974954490b6p-37);
  p = __builtin_fma(p, z, 0x1.1eed8eff336c1p-29);
  p = __builtin_fma(p, z, -0x1.27e4fb7789d98p-22);
  p = __builtin_fma(p, z, 0x1.a01a01a01a019p-16);
  p = __builtin_fma(p, z, -0x1.6c16c16c16c17p-10);
  p = __builtin_fma(p, z, 0x1.5555555555555p-5);
  return p;
}
// sin(h + l), cos(h + l) for |h| <= pi/4 + small, |l| <= ulp(h).
__device__ __forceinline__ double sin_kernel(double h, double l) {
  const double z = h * h;
  const double s = (h * z) * sin_poly(z);
  return h + __builtin_fma(l, __builtin_fma(-0.5, z, 1.0), s);
}
__device__ __forceinline__ double cos_kernel(double h, double l) {
  const double z = h * h;
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  const double e = (1.0 - w) - hz;  // exact: 1 - w is exact for w in [1/2, 1]
  return w + (__builtin_fma(z * z, cos_poly(z), e) - h * l);
}

// Round-to-nearest sin / cos of x in [0, 8] (the interval functions reduce
// their arguments to [0, 2 pi] first; larger |x| is handled by them).
__device__ __forceinline__ void sincos_rn(double x, double* sn, double* cs) {
  const double k = __builtin_rint(x * trig::TWO_OVER_PI);
  const double r1 = __builtin_fma(-k, trig::P1, x);  // exact for 0 <= x < 8
  const double th = k * trig::P2;
  const double tl = __builtin_fma(k, trig::P2, -th);
  const double s = r1 - th;  // two-sum of r1 and -th
  const double bb = s - r1;
  const double err = (r1 - (s - bb)) + (-th - bb);
  const double lo = err - __builtin_fma(k, trig::P3, tl);
  const double h = s + lo;
  const double l = lo - (h - s);
  const double sk = sin_kernel(h, l), ck = cos_kernel(h, l);
  const int q = (int)k & 3;
  *sn = q == 0 ? sk : q == 1 ? ck : q == 2 ? -sk : -ck;
  *cs = q == 0 ? ck : q == 1 ? -sk : q == 2 ? -ck : sk;
}
__device__ __forceinline__ double sin_rn(double x) {
  double s, c;
  sincos_rn(x, &s, &c);
  return s;
}
__device__ __forceinline__ double cos_rn(double x) {
  double s, c;
  sincos_rn(x, &s, &c);
  return c;
}

// acos from the arcsine kernel: |x| <= 1/2: pi/2 - asin(x); x > 1/2:
// 2 asin(sqrt((1 - x) / 2)); x < -1/2: pi - 2 asin(sqrt((1 + x) / 2)).
__device__ __forceinline__ double acos_rn(double x) {
  const double ax = __builtin_fabs(x);
  if (ax <= 0.5) {
    const double z = x * x;
    const double q = (x * z) * asin_poly_eval(z);
    return trig::P1 - (x - (trig::P2 - q));
  }
  const double t = 0.5 - 0.5 * ax;
  const double s = __builtin_sqrt(t);
  const double c = t > 0.0 ? __builtin_fma(-s, s, t) / (s + s) : 0.0;
  const double w = __builtin_fma(s * t, asin_poly_eval(t), c);
  if (x > 0.0) return 2.0 * s + 2.0 * w;
  return 2.0 * trig::P1 - ((2.0 * s + 2.0 * w) - 2.0 * trig::P2);
}

__device__ __forceinline__ double steps_toward(double y, double toward, int n) {
  for (int i = 0; i < n; ++i) y = next_toward(y, toward);
  return y;
}
// interval.cuh:376-418 with our functions
__device__ __forceinline__ double dsin_rd(double x) {
  double y = steps_toward(sin_rn(x), -2.0, SC_STEPS);
  return y < -1.0 ? -1.0 : y;
}
__device__ __forceinline__ double dsin_ru(double x) {
  double y = steps_toward(sin_rn(x), 2.0, SC_STEPS);
  return y > 1.0 ? 1.0 : y;
}
__device__ __forceinline__ double dcos_rd(double x) {
  double y = steps_toward(cos_rn(x), -2.0, SC_STEPS);
  return y < -1.0 ? -1.0 : y;
}
__device__ __forceinline__ double dcos_ru(double x) {
  double y = steps_toward(cos_rn(x), 2.0, SC_STEPS);
  return y > 1.0 ? 1.0 : y;
}
__device__ __forceinline__ double dacos_ru(double x) {
  double y = steps_toward(acos_rn(x), 4.0, SC_STEPS);
  return y > trig::ub_pi ? trig::ub_pi : y;
}
__device__ __forceinline__ double dacos_rd(double x) {
  double y = steps_toward(acos_rn(x), -1.0, SC_STEPS);
  return y < 0.0 ? 0.0 : y;
}

// interval.cuh:440-478
__device__ inline IV sin_in_range(IV x) {
  using namespace trig;
  if (x.ub <= ub_pi_two) {
    if (x.ub >= lb_pi_two) {
      const double lb1 = dsin_rd(x.lb), lb2 = dsin_rd(ub_pi_two);
      return IV(lb1 < lb2 ? lb1 : lb2, 1.0);
    }
    return IV(dsin_rd(x.lb), dsin_ru(x.ub));
  }
  if (x.ub <= ub_3pi_two) {
    if (x.lb >= ub_pi_two) {
      const double ub = dsin_ru(x.lb);
      const double lb = x.ub >= lb_3pi_two ? -1.0 : dsin_rd(x.ub);
      return IV(lb, ub);
    }
    const double lb1 = dsin_rd(x.lb), lb2 = dsin_rd(x.ub);
    return IV(lb1 < lb2 ? lb1 : lb2, 1.0);
  }
  if (x.lb <= ub_pi_two) return IV(-1.0, 1.0);
  if (x.lb <= ub_3pi_two) {
    const double ub1 = dsin_ru(x.lb), ub2 = dsin_ru(x.ub);
    return IV(-1.0, ub1 > ub2 ? ub1 : ub2);
  }
  return IV(dsin_rd(x.lb), dsin_ru(x.ub));
}
// interval.cuh:480-499
__device__ inline IV cos_in_range(IV x) {
  if (x.ub <= trig::ub_pi) return IV(x.ub >= trig::lb_pi ? -1.0 : dcos_rd(x.ub), dcos_ru(x.lb));
  if (x.lb <= trig::ub_pi) {
    const double ub1 = dcos_ru(x.lb), ub2 = dcos_ru(x.ub);
    return IV(-1.0, ub1 > ub2 ? ub1 : ub2);
  }
  return IV(dcos_rd(x.lb), dcos_ru(x.ub));
}

// interval.cuh:501-576: reduce a non-negative interval to [0, 2 pi] by
// directed multiples of 2 pi (one period, or two adjacent ones joined).
template <bool COS>
__device__ inline IV trig_non_negative(IV x) {
  using namespace trig;
  if (x.lb >= 0.0 && x.ub <= ub_two_pi) return COS ? cos_in_range(x) : sin_in_range(x);
  const double li = __builtin_trunc(div_rd(x.lb, ub_two_pi));
  const double ui = __builtin_trunc(div_ru(x.ub, lb_two_pi));
  const double diff = ui - li;
  if (diff > 1.0) return IV(-1.0, 1.0);
  if (diff == 0.0) {
    double lr = add_rd(x.lb, -mul_ru(li, ub_two_pi));
    double ur = add_ru(x.ub, -mul_rd(li, lb_two_pi));
    lr = lr < 0.0 ? 0.0 : (lr > ub_two_pi ? ub_two_pi : lr);
    ur = ur < 0.0 ? 0.0 : (ur > ub_two_pi ? ub_two_pi : ur);
    return COS ? cos_in_range(IV(lr, ur)) : sin_in_range(IV(lr, ur));
  }
  double lr = add_rd(x.lb, -mul_ru(li, ub_two_pi));
  if (lr < 0.0) return IV(-1.0, 1.0);
  if (lr > ub_two_pi) lr = ub_two_pi;
  double ur = add_ru(x.ub, -mul_rd(ui, lb_two_pi));
  if (ur > ub_two_pi) return IV(-1.0, 1.0);
  if (ur < 0.0) ur = 0.0;
  IV a = COS ? cos_in_range(IV(lr, ub_two_pi)) : sin_in_range(IV(lr, ub_two_pi));
  IV b = COS ? cos_in_range(IV(0.0, ur)) : sin_in_range(IV(0.0, ur));
  return a.do_join(b);
}

__device__ inline IV neg(IV x) { return IV(-x.ub, -x.lb); }

// interval.cuh:578-589
__device__ inline IV sin_iv(IV x) {
  if (x.lb >= 0.0) return trig_non_negative<false>(x);
  if (x.ub <= 0.0) return neg(trig_non_negative<false>(neg(x)));
  IV p = trig_non_negative<false>(IV(0.0, x.ub));
  return p.do_join(neg(trig_non_negative<false>(IV(0.0, -x.lb))));
}
// interval.cuh:591-601
__device__ inline IV cos_iv(IV x) {
  if (x.lb >= 0.0) return trig_non_negative<true>(x);
  if (x.ub <= 0.0) return trig_non_negative<true>(neg(x));
  const double nl = -x.lb;
  return trig_non_negative<true>(IV(0.0, nl < x.ub ? x.ub : nl));
}
// interval.cuh:603-610 (trap on an argument outside [-1, 1])
__device__ inline IV acos_iv(IV x, int& trap) {
  if (x.lb < -1.0 || x.ub > 1.0) {
    if (!trap) trap = T_ASIN;
    return IV(0.0, 0.0);
  }
  return IV(dacos_rd(x.ub), dacos_ru(x.lb));
}

// interval.cuh:330-350 (trap on a possibly negative argument)
__device__ inline IV sqrt_iv(IV x, int& trap) {
  if (x.lb < 0.0) {
    if (!trap) trap = 5;  // CPL_TRAP_SQRT_DOMAIN
    return IV(0.0, 0.0);
  }
  return IV(sqrt_rd(x.lb), sqrt_ru(x.ub));
}
// interval.cuh:227-235.  For a negative interval the reference returns
// {rcp_rd(lb), rcp_ru(ub)}, which is inverted (empty); the enclosure
// [rd(1/ub), ru(1/lb)] is returned here instead.
__device__ inline IV reciprocal_iv(IV x) {
  if (x.lb > 0.0) return IV(div_rd(1.0, x.ub), div_ru(1.0, x.lb));
  if (x.ub < 0.0) return IV(div_rd(1.0, x.ub), div_ru(1.0, x.lb));
  const double inf = __builtin_huge_val();
  return IV(-inf, inf);
}
// interval.cuh:261-275
__device__ inline IV div_scalar(IV x, double n) {
  if (n > 0.0) return IV(div_rd(x.lb, n), div_ru(x.ub, n));
  return IV(div_rd(x.ub, n), div_ru(x.lb, n));
}

}  // namespace cpl
