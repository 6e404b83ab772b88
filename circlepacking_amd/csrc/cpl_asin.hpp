// The prover's own round-to-nearest arcsine and the reference's widened
// enclosures dasin_rd / dasin_ru (algcuda/cuda-utils/cuh/include/algcuda/interval.cuh:420-438).
//
// The reference widens CUDA's ::asin by five ulps.  CUDA's implementation is
// not reproduced: asin_rn is an independent design (argument reduction to
// [0, 1/2] plus a degree-12 Chebyshev polynomial in x^2; coefficients from
// gen_asin_poly.py), written only with correctly rounded IEEE operations in
// round-to-nearest (+, -, *, fma, sqrt) so its bits are reproducible on any
// IEEE machine.  Measured max error 1.50 ulp over dense sweeps of [0, 1]
// (near x = 0.506, where the rounding error of sqrt(t) is not corrected); the
// five-ulp widening encloses the true arcsine with room for a step across a
// binade edge (5 > 2 x 1.5 + 1).
#pragma once

#include "cpl_iv.hpp"

namespace cpl {

// Horner steps as fma_sc (coefficient in scalar registers, see cpl_round.hpp).
__device__ __forceinline__ double asin_poly_eval(double z) {
  double p = 0x1.d72b2bc8155f8p-6;
  p = fma_sc(p, z, -0x1.e6aaa8a0a04ccp-7);
  p = fma_sc(p, z, 0x1.1d189408314eep-6);
  p = fma_sc(p, z, 0x1.65a9c4dfcf8b2p-8);
  p = fma_sc(p, z, 0x1.52420b04b37bep-7);
  p = fma_sc(p, z, 0x1.782651caa6547p-7);
  p = fma_sc(p, z, 0x1.c9cf07674736ap-7);
  p = fma_sc(p, z, 0x1.1c4d35cf95421p-6);
  p = fma_sc(p, z, 0x1.6e8bb1c8209a2p-6);
  p = fma_sc(p, z, 0x1.f1c71c1db0623p-6);
  p = fma_sc(p, z, 0x1.6db6db6e31f13p-5);
  p = fma_sc(p, z, 0x1.3333333332ecap-4);
  p = fma_sc(p, z, 0x1.5555555555556p-3);
  return p;
}

// Round-to-nearest square root of t in [0, 1/4]: the hardware reciprocal
// square root and the Newton sequence the compiler emits for sqrt(double)
// (g = t y, h = y / 2, two half-steps, two residual corrections), without its
// scaling of arguments below 2^-767 and its zero / infinity fix-up, neither of
// which t in [2^-54, 1/4] needs (t = +-0 returns t, t < 0 NaN, as sqrt does).
// Bit-identical to the compiler's sqrt there, which is correctly rounded.
__device__ __forceinline__ double sqrt_quarter(double t) {
  const double y = __builtin_amdgcn_rsq(t);
  double g = t * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  double d = __builtin_fma(-g, g, t);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, t);
  g = __builtin_fma(d, h, g);
  return t > 0.0 ? g : (t == 0.0 ? t : __builtin_nan(""));
}

// Both ranges evaluate the one polynomial (in x^2 below 1/2, in t = (1 - |x|)/2
// above), so a wave whose lanes fall in both ranges evaluates it once.
struct AsinArg {
  double ax, t, z;
  bool small;
};
__device__ __forceinline__ AsinArg asin_arg(double x) {
  AsinArg a;
  a.ax = __builtin_fabs(x);
  a.small = a.ax <= 0.5;
  a.t = 0.5 - 0.5 * a.ax;  // exact for ax in [1/2, 1]
  a.z = a.small ? a.ax * a.ax : a.t;
  return a;
}
__device__ __forceinline__ double asin_finish(double x, const AsinArg& a, double p) {
  double r;
  if (a.small) {
    r = __builtin_fma(a.ax * a.z, p, a.ax);
  } else {
    // asin(sqrt t) ~= s + w; the correction of s's own rounding error, (t - s^2) / 2s,
    // would cost an IEEE division for at most 0.4 ulp of the result
    double s = sqrt_quarter(a.t);
    double w = (s * a.t) * p;
    r = 0x1.921fb54442d18p+0 - ((s + s) + (w + w - 0x1.1a62633145c07p-54));
  }
  return __builtin_copysign(r, x);
}
__device__ __forceinline__ double asin_rn(double x) {
  const AsinArg a = asin_arg(x);
  return asin_finish(x, a, asin_poly_eval(a.z));
}
// Two arcsines with their polynomials evaluated side by side (two independent
// Horner chains the scheduler interleaves); each result is asin_rn's.
__device__ __forceinline__ void asin_rn_pair(double x1, double x2, double& r1, double& r2) {
  const AsinArg a1 = asin_arg(x1), a2 = asin_arg(x2);
  const double p1 = asin_poly_eval(a1.z), p2 = asin_poly_eval(a2.z);
  r1 = asin_finish(x1, a1, p1);
  r2 = asin_finish(x2, a2, p2);
}

// nextafter(y, toward) for the arguments dasin uses (toward = +-2.0).
__device__ __forceinline__ double next_toward(double y, double toward) {
  if (__builtin_isnan(y)) return y;
  if (y == toward) return toward;
  if (y == 0.0) return __builtin_copysign(0x1p-1074, toward);
  uint64_t b = dbits(y);
  b = ((y < toward) == (y > 0.0)) ? b + 1 : b - 1;
  return dfrom(b);
}

// Five ulp steps outward: a direct bit step when no zero crossing is possible.
__device__ __forceinline__ double step5(double y, bool up) {
  if (__builtin_fabs(y) >= 0x1p-1000 && __builtin_fabs(y) < 1.9) {
    uint64_t b = dbits(y);
    bool away = up == (y > 0.0);
    return dfrom(away ? b + 5 : b - 5);
  }
  double t = up ? 2.0 : -2.0;
  for (int i = 0; i < 5; ++i) y = next_toward(y, t);
  return y;
}

// For x = +-0 both widened bounds are +0: five steps from asin(+-0) = +-0
// cross zero and the sign clamps (x >= 0 for the lower, x <= 0 for the upper
// bound) bring them back.  Zero arguments are common (clamped quotients of the
// collision angles), and taking them early keeps the stepping loop of step5,
// which a wave runs whenever one of its lanes needs it, off their path.
//
// Wave-uniform fast path: when every lane's argument lies in [2^-1000, 1] (the
// clamped quotients of the collision angles), asin_rn(x) lies in
// [2^-1000, pi/2], so step5 is the direct bit step (5 ulps down for the lower,
// up for the upper bound), the zero and sign clamps cannot apply and only the
// upper bound's pi/2 clamp can: the same bits as the general path.
__device__ __forceinline__ bool asin_unit_arg(double x) { return x >= 0x1p-1000 && x <= 1.0; }
__device__ __forceinline__ double dasin_rd(double x) {
  if (all_lanes(asin_unit_arg(x))) return dfrom(dbits(asin_rn(x)) - 5);
  if (x == 0.0) return 0.0;
  double y = step5(asin_rn(x), false);
  if (y < -1.5707963267948968) y = -1.5707963267948968;
  if (x >= 0.0 && y < 0.0) y = 0.0;
  return y;
}
__device__ __forceinline__ double dasin_ru(double x) {
  if (all_lanes(asin_unit_arg(x))) {
    const double y = dfrom(dbits(asin_rn(x)) + 5);
    return y > 1.5707963267948968 ? 1.5707963267948968 : y;
  }
  if (x == 0.0) return 0.0;
  double y = step5(asin_rn(x), true);
  if (y > 1.5707963267948968) y = 1.5707963267948968;
  if (x <= 0.0 && y > 0.0) y = 0.0;
  return y;
}
// Two widened arcsines (upper bound if UP1 / UP2) of independent arguments:
// when every lane's arguments lie in [2^-1000, 1] both polynomials run side
// by side; the same bits as dasin_rd / dasin_ru of each.
template <bool UP1, bool UP2>
__device__ __forceinline__ void dasin_pair(double x1, double x2, double& y1, double& y2) {
  if (all_lanes(asin_unit_arg(x1) && asin_unit_arg(x2))) {
    double a1, a2;
    asin_rn_pair(x1, x2, a1, a2);
    y1 = dfrom(UP1 ? dbits(a1) + 5 : dbits(a1) - 5);
    y2 = dfrom(UP2 ? dbits(a2) + 5 : dbits(a2) - 5);
    if (UP1 && y1 > 1.5707963267948968) y1 = 1.5707963267948968;
    if (UP2 && y2 > 1.5707963267948968) y2 = 1.5707963267948968;
    return;
  }
  y1 = UP1 ? dasin_ru(x1) : dasin_rd(x1);
  y2 = UP2 ? dasin_ru(x2) : dasin_rd(x2);
}

}  // namespace cpl

// ============================================================================
// FP32 arcsine for the outward-rounded prefilter.  Same reduction as asin_rn
// with the degree-5 interpolant rounded to float (gen_asin_poly.py --fp32);
// its error against the exact arcsine stays below 2 ulps (measured by the
// parity tests over 2.4e6 arguments), and the bounds step ASIN32_STEPS = 6
// ulps outward (more than twice the error plus one, so a step across a binade
// edge still leaves room).
// ============================================================================
namespace cpl {

constexpr int ASIN32_STEPS = 6;

__device__ __forceinline__ float asin32_poly(float z) {
  float p = 0x1.13fed4p-5f;
  p = __builtin_fmaf(p, z, 0x1.18f91ep-6f);
  p = __builtin_fmaf(p, z, 0x1.fd8da2p-6f);
  p = __builtin_fmaf(p, z, 0x1.6d5bbap-5f);
  p = __builtin_fmaf(p, z, 0x1.33343p-4f);
  p = __builtin_fmaf(p, z, 0x1.555554p-3f);
  return p;
}

__device__ __forceinline__ float asin32_rn(float x) {
  float ax = __builtin_fabsf(x);
  const bool small = ax <= 0.5f;
  float t = 0.5f - 0.5f * ax;
  float z = small ? ax * ax : t;
  float p = asin32_poly(z);
  float r;
  if (small) {
    r = __builtin_fmaf(ax * z, p, ax);
  } else {
    float s = __builtin_sqrtf(t);
    float w = __builtin_fmaf(s * t, p, 0.0f);
    r = 0x1.921fb6p+0f - ((s + s) + (w + w + 0x1.777a5cp-25f));
  }
  return __builtin_copysignf(r, x);
}

__device__ __forceinline__ float next_toward_f(float y, float toward) {
  if (y != y) return y;
  if (y == toward) return toward;
  if (y == 0.0f) return __builtin_copysignf(0x1p-149f, toward);
  int b = __float_as_int(y);
  b = ((y < toward) == (y > 0.0f)) ? b + 1 : b - 1;
  return __int_as_float(b);
}

__device__ __forceinline__ float steps_out_f(float y, bool up) {
  if (__builtin_fabsf(y) >= 0x1p-100f && __builtin_fabsf(y) < 1.9f) {
    int b = __float_as_int(y);
    bool away = up == (y > 0.0f);
    return __int_as_float(away ? b + ASIN32_STEPS : b - ASIN32_STEPS);
  }
  float t = up ? 2.0f : -2.0f;
  for (int i = 0; i < ASIN32_STEPS; ++i) y = next_toward_f(y, t);
  return y;
}

// Clamps of interval.cuh:426-437 with pi/2 rounded outward to FP32.
__device__ __forceinline__ float dasin_rd(float x) {
  float y = steps_out_f(asin32_rn(x), false);
  if (y < -0x1.921fb6p+0f) y = -0x1.921fb6p+0f;
  if (x >= 0.0f && y < 0.0f) y = 0.0f;
  return y;
}
__device__ __forceinline__ float dasin_ru(float x) {
  float y = steps_out_f(asin32_rn(x), true);
  if (y > 0x1.921fb6p+0f) y = 0x1.921fb6p+0f;
  if (x <= 0.0f && y > 0.0f) y = 0.0f;
  return y;
}

}  // namespace cpl
