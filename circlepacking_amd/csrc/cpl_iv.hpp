// Device interval type for the lane-packing prover.
//
// Mirrors algcuda::Interval<NumType> / Uncertain<bool>
// (algcuda/cuda-utils/cpp/include/algcuda/interval.hpp:32-242) and the device
// arithmetic of algcuda/cuda-utils/cuh/include/algcuda/interval.cuh:30-355,
// on top of the MODE-register primitives of cpl_round.hpp.  Same POD layout
// {lb, ub}; same operator set and semantics, so the criteria read like the
// reference's.  IV = IVT<double> is the certifying type; IVT<float> is the
// outward-rounded prefilter type.
#pragma once

#include "cpl_round.hpp"

namespace cpl {

// True when c holds on every active lane of the wave (a wave-uniform branch
// condition for the sign-case fast paths below; inactive lanes do not vote).
__device__ __forceinline__ bool all_lanes(bool c) { return __ballot(!c) == 0; }

struct UB {
  bool lb, ub;  // lb: definitely, ub: possibly (interval.hpp:32-82)
  __device__ UB operator!() const { return UB{!ub, !lb}; }
};
__device__ __forceinline__ bool definitely(UB b) { return b.lb; }
__device__ __forceinline__ bool possibly(UB b) { return b.ub; }

template <class T>
struct IVT {
  T lb, ub;
  IVT() = default;
  __host__ __device__ constexpr IVT(T l, T u) : lb(l), ub(u) {}

  __device__ bool empty() const { return lb > ub; }
  __device__ void tighten_lb(T v) { if (lb < v) lb = v; }
  __device__ void tighten_ub(T v) { if (v < ub) ub = v; }
  __device__ IVT& do_join(const IVT& o) {
    if (o.lb < lb) lb = o.lb;
    if (o.ub > ub) ub = o.ub;
    return *this;
  }
  __device__ IVT intersect(const IVT& o) const { return IVT(lb < o.lb ? o.lb : lb, ub > o.ub ? o.ub : ub); }
  __device__ IVT abs() const {  // interval.hpp:216-232
    T l = T(0), u = T(0);
    if (ub > 0) u = ub; else l = -ub;
    if (lb < 0) u = u < -lb ? -lb : u; else l = lb;
    return IVT(l, u);
  }
  __device__ IVT square() const {  // interval.cuh:352-355
    IVT r;
    if (all_lanes(lb >= T(0) && ub > T(0))) {
      // abs() is the identity on such an interval (a -0 lower bound included)
      mul_pair(lb, lb, ub, ub, r.lb, r.ub);
      return r;
    }
    IVT a = abs();
    mul_pair(a.lb, a.lb, a.ub, a.ub, r.lb, r.ub);
    return r;
  }
};
using IV = IVT<double>;
using IVf = IVT<float>;

template <class T>
__device__ __forceinline__ IVT<T> operator+(const IVT<T>& a, const IVT<T>& b) {
  IVT<T> r;
  add_pair(a.lb, b.lb, a.ub, b.ub, r.lb, r.ub);
  return r;
}
template <class T>
__device__ __forceinline__ IVT<T> operator+(const IVT<T>& a, T b) { return a + IVT<T>(b, b); }
template <class T>
__device__ __forceinline__ IVT<T> operator+(T a, const IVT<T>& b) { return b + IVT<T>(a, a); }
template <class T>
__device__ __forceinline__ IVT<T> operator-(const IVT<T>& a, const IVT<T>& b) {
  IVT<T> r;
  sub_pair(a.lb, b.ub, a.ub, b.lb, r.lb, r.ub);
  return r;
}
template <class T>
__device__ __forceinline__ IVT<T> operator-(const IVT<T>& a, T b) { return a - IVT<T>(b, b); }
template <class T>
__device__ __forceinline__ IVT<T> operator-(T a, const IVT<T>& b) { return IVT<T>(a, a) - b; }

// Interval x scalar (interval.cuh:185-197); for n >= 0 the upper bound is
// rounded up (the reference rounds it down; identical for the power-of-two
// factors 0.25, 0.5, 2 of every call site on this path).
template <class T>
__device__ __forceinline__ IVT<T> operator*(const IVT<T>& a, T n) {
  IVT<T> r;
  if (n >= 0) mul_pair(a.lb, n, a.ub, n, r.lb, r.ub);
  else mul_pair(a.ub, n, a.lb, n, r.lb, r.ub);
  return r;
}
template <class T>
__device__ __forceinline__ IVT<T> operator*(T n, const IVT<T>& a) { return a * n; }

// Interval x interval, 8-product form (interval.cuh:171-182).
template <class T>
__device__ __forceinline__ IVT<T> operator*(const IVT<T>& a, const IVT<T>& b) {
  T l1, l2, l3, l4, u1, u2, u3, u4;
  mul_pair(a.lb, b.lb, a.lb, b.lb, l1, u1);
  mul_pair(a.lb, b.ub, a.lb, b.ub, l2, u2);
  mul_pair(a.ub, b.lb, a.ub, b.lb, l3, u3);
  mul_pair(a.ub, b.ub, a.ub, b.ub, l4, u4);
  T ll1 = l1 < l2 ? l1 : l2, ll2 = l3 < l4 ? l3 : l4;
  T uu1 = u1 > u2 ? u1 : u2, uu2 = u3 > u4 ? u3 : u4;
  return IVT<T>(ll1 < ll2 ? ll1 : ll2, uu1 > uu2 ? uu1 : uu2);
}

// Interval / interval with the sign cases of interval.cuh:277-316.  The six
// cases differ only in which endpoints are divided, so the operands are
// selected branch-free and one directed pair is computed; a divisor that
// contains zero gives (-inf, +inf) as in the reference.
template <class T>
__device__ __forceinline__ IVT<T> operator/(const IVT<T>& a, const IVT<T>& o) {
  if (all_lanes(o.lb > T(0) && a.lb >= T(0))) {
    // the case the sign selection below picks for a >= 0 over a positive divisor
    IVT<T> r;
    div_pair(a.lb, o.ub, a.ub, o.lb, r.lb, r.ub);
    return r;
  }
  const bool pos = o.lb > 0, neg = o.ub < 0;
  const bool anneg = a.lb >= 0, anpos = a.ub <= 0;  // a >= 0, a <= 0
  T nl, dl, nu, du;
  if (pos) {
    nl = a.lb;
    dl = anneg ? o.ub : o.lb;
    nu = a.ub;
    du = (anneg || !anpos) ? o.lb : o.ub;
  } else {
    nl = a.ub;
    dl = (anneg || !anpos) ? o.ub : o.lb;
    nu = a.lb;
    du = anneg ? o.lb : o.ub;
  }
  if (!pos && !neg) { nl = nu = T(0); dl = du = T(1); }
  IVT<T> r;
  div_pair(nl, dl, nu, du, r.lb, r.ub);
  if (!pos && !neg) {
    r.ub = (T)__builtin_huge_val();
    r.lb = -r.ub;
  }
  return r;
}

template <class T>
__device__ __forceinline__ UB operator<(const IVT<T>& a, const IVT<T>& b) { return UB{a.ub < b.lb, a.lb < b.ub}; }
template <class T>
__device__ __forceinline__ UB operator<=(const IVT<T>& a, const IVT<T>& b) { return UB{a.ub <= b.lb, a.lb <= b.ub}; }
template <class T>
__device__ __forceinline__ UB operator>(const IVT<T>& a, const IVT<T>& b) { return b < a; }
template <class T>
__device__ __forceinline__ UB operator>=(const IVT<T>& a, const IVT<T>& b) { return b <= a; }

// get_subinterval (subintervals.hpp:31-42) in round-to-nearest with the
// multiply-add fused, as nvcc compiles lb + diff*i.
__device__ __forceinline__ IV get_subinterval(IV iv, int i, int n) {
  double diff = (iv.ub - iv.lb) / (double)n;
  double l = __builtin_fma(diff, (double)i, iv.lb);
  double u = (i + 1 == n) ? iv.ub : __builtin_fma(diff, (double)(i + 1), iv.lb);
  return IV(l, u);
}

// Outward FP64 -> FP32 interval.
__device__ __forceinline__ IVf to_f32(const IV& x) { return IVf(f32_down(x.lb), f32_up(x.ub)); }

}  // namespace cpl
