// Directed-rounding FP64 primitives for gfx950 (CDNA4).
//
// CUDA has per-instruction rounding (__dadd_rd, __dmul_ru, ...).  CDNA has a
// per-wave rounding MODE instead: MODE[3:2] is the FP64/FP16 rounding field
// (0 nearest-even, 1 toward +inf, 2 toward -inf, 3 toward zero), MODE[1:0] the
// FP32 one.  Every directed operation here is one asm statement that
//     s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), <RD|RU>
//     <the IEEE instruction(s)>
//     s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0
// so the wave is back in round-to-nearest whenever compiler-generated code
// runs.  Because each statement is self-contained and a pure function of its
// operands, the compiler may schedule, hoist or merge them freely.
//
// Hazards: two s_setreg of the same hardware register need >= 2 wait states
// between them (the rule the compiler's hazard recognizer applies to its own
// s_setreg sequences).  Every statement therefore starts with s_nop 1 (a
// preceding statement may have ended with an s_setreg) and pads a single
// instruction between its own s_setreg writes with s_nop 0.
//
// add, sub and mul are single IEEE instructions, correctly rounded in the
// selected mode.  Division and square root are multi-instruction sequences
// whose directed-mode results are NOT correctly rounded, so they are computed
// as: a faithful quotient q (div_faithful: one of the two doubles around a/b),
// then one fused multiply-add pair executed in the target mode,
//     r = fma(-q, b, a)          (the exact remainder, rounded toward the target)
//     result = fma(r, y, q)      (y a power of two, |y b| < 1)
// whose directed rounding lands on q or its neighbour exactly as the rounded
// true quotient does (Markstein-style correction; see div_fast_ok below).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpl {

#define CPL_MODE_RD "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 2\n\t"
#define CPL_MODE_RU "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 1\n\t"
#define CPL_MODE_RN "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0\n\t"

// ---- one instruction in one mode -------------------------------------------
#define CPL_DIR_BINOP(NAME, MODE, INSN)                                                      \
  __device__ __forceinline__ double NAME(double a, double b) {                              \
    double r;                                                                               \
    asm("s_nop 1\n\t" MODE INSN " %0, %1, %2\n\ts_nop 0\n\t" CPL_MODE_RN                    \
        : "=v"(r)                                                                           \
        : "v"(a), "v"(b));                                                                  \
    return r;                                                                               \
  }
CPL_DIR_BINOP(add_rd, CPL_MODE_RD, "v_add_f64")
CPL_DIR_BINOP(add_ru, CPL_MODE_RU, "v_add_f64")
CPL_DIR_BINOP(mul_rd, CPL_MODE_RD, "v_mul_f64")
CPL_DIR_BINOP(mul_ru, CPL_MODE_RU, "v_mul_f64")
#undef CPL_DIR_BINOP

__device__ __forceinline__ double sub_rd(double a, double b) {
  double r;
  asm("s_nop 1\n\t" CPL_MODE_RD "v_add_f64 %0, %1, -%2\n\ts_nop 0\n\t" CPL_MODE_RN : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double sub_ru(double a, double b) {
  double r;
  asm("s_nop 1\n\t" CPL_MODE_RU "v_add_f64 %0, %1, -%2\n\ts_nop 0\n\t" CPL_MODE_RN : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// ---- interval-shaped pairs: lb in RD, ub in RU, one statement ---------------
// [rd(al + bl), ru(au + bu)]
__device__ __forceinline__ void add_pair(double al, double bl, double au, double bu, double& l, double& u) {
  asm("s_nop 1\n\t" CPL_MODE_RD "v_add_f64 %0, %2, %3\n\ts_nop 0\n\t" CPL_MODE_RU
      "v_add_f64 %1, %4, %5\n\ts_nop 0\n\t" CPL_MODE_RN
      : "=&v"(l), "=v"(u)
      : "v"(al), "v"(bl), "v"(au), "v"(bu));
}
// [rd(al - bu), ru(au - bl)]
__device__ __forceinline__ void sub_pair(double al, double bu, double au, double bl, double& l, double& u) {
  asm("s_nop 1\n\t" CPL_MODE_RD "v_add_f64 %0, %2, -%3\n\ts_nop 0\n\t" CPL_MODE_RU
      "v_add_f64 %1, %4, -%5\n\ts_nop 0\n\t" CPL_MODE_RN
      : "=&v"(l), "=v"(u)
      : "v"(al), "v"(bu), "v"(au), "v"(bl));
}
// [rd(al * bl), ru(au * bu)]
__device__ __forceinline__ void mul_pair(double al, double bl, double au, double bu, double& l, double& u) {
  asm("s_nop 1\n\t" CPL_MODE_RD "v_mul_f64 %0, %2, %3\n\ts_nop 0\n\t" CPL_MODE_RU
      "v_mul_f64 %1, %4, %5\n\ts_nop 0\n\t" CPL_MODE_RN
      : "=&v"(l), "=v"(u)
      : "v"(al), "v"(bl), "v"(au), "v"(bu));
}

// ---- several pairs in one statement -----------------------------------------
// Each mode switch costs the wave about 30 cycles and four scalar instructions
// (s_nop, s_setreg, s_nop, s_setreg around the operation); independent
// directed operations of the same mode share one switch here.  Two or more
// instructions always separate consecutive s_setreg writes.
//
// [rd(a1 b1), ru(c1 d1)] and [rd(a2 b2), ru(c2 d2)]
__device__ __forceinline__ void mul_pair2(double a1, double b1, double c1, double d1, double a2, double b2, double c2,
                                          double d2, double& l1, double& u1, double& l2, double& u2) {
  asm("s_nop 1\n\t" CPL_MODE_RD "v_mul_f64 %0, %4, %5\n\tv_mul_f64 %2, %8, %9\n\t" CPL_MODE_RU
      "v_mul_f64 %1, %6, %7\n\tv_mul_f64 %3, %10, %11\n\ts_nop 0\n\t" CPL_MODE_RN
      : "=&v"(l1), "=&v"(u1), "=&v"(l2), "=v"(u2)
      : "v"(a1), "v"(b1), "v"(c1), "v"(d1), "v"(a2), "v"(b2), "v"(c2), "v"(d2));
}
// With h >= 0 (the factor 1/2 of the collision-angle construction):
//   lx = [rd(h rd(tl + al)), ru(h ru(tu + au))]   (h * (t + a))
//   sx = [rd(h rd(tl - au)), ru(h ru(tu - al))]   (h * (t - a))
// the same operations, in the same order per result, as an interval sum and
// difference followed by interval x scalar.
__device__ __forceinline__ void half_sum_diff(double h, double tl, double tu, double al, double au, double& lxl,
                                              double& lxu, double& sxl, double& sxu) {
  double s0, s1;
  asm("s_nop 1\n\t" CPL_MODE_RD "v_add_f64 %4, %6, %8\n\tv_add_f64 %5, %6, -%9\n\t"
      "v_mul_f64 %0, %4, %10\n\tv_mul_f64 %2, %5, %10\n\t" CPL_MODE_RU
      "v_add_f64 %4, %7, %9\n\tv_add_f64 %5, %7, -%8\n\t"
      "v_mul_f64 %1, %4, %10\n\tv_mul_f64 %3, %5, %10\n\ts_nop 0\n\t" CPL_MODE_RN
      : "=&v"(lxl), "=&v"(lxu), "=&v"(sxl), "=&v"(sxu), "=&v"(s0), "=&v"(s1)
      : "v"(tl), "v"(tu), "v"(al), "v"(au), "v"(h));
}

// ---- fma with a scalar-register addend --------------------------------------
// fma(a, b, c) in round-to-nearest (the ambient mode) as the three-operand
// v_fma_f64 with c in an SGPR pair.  For a constant c the compiler otherwise
// picks the two-address v_fmac_f64, whose accumulator must be a VGPR pair:
// two v_mov per Horner step (the 64-bit constant cannot be an operand literal).
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}

// ---- bit helpers ------------------------------------------------------------
__device__ __forceinline__ uint64_t dbits(double d) { return __double_as_longlong(d); }
__device__ __forceinline__ double dfrom(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ int dexp_field(double d) { return (int)((dbits(d) >> 52) & 0x7FF); }
// Atomic_ordered_double's order-preserving map (atomic_float.hpp:56-66): an
// unsigned comparison of the images orders the doubles, so atomicMin on the
// u64 is a minimum over doubles.
__device__ __forceinline__ unsigned long long ordered_of(double v) {
  uint64_t b = dbits(v);
  uint64_t mask = (0 - (b >> 63)) | (1ull << 63);
  return b ^ mask;
}
__device__ __forceinline__ double double_of_ordered(unsigned long long v) {
  uint64_t mask = ((v >> 63) - 1) | (1ull << 63);
  return dfrom(v ^ mask);
}
// 2^k for -1022 <= k <= 1023 (normal powers of two only)
__device__ __forceinline__ double pow2(int k) { return dfrom((uint64_t)(k + 1023) << 52); }

// ---- division ---------------------------------------------------------------
// Correction step of a directed division, both fmas in the target mode:
//   r = fma(-q, b, a), res = fma(r, y, q).
// With b > 0, q = RN(a/b) and y = 2^-(e_b+1) (so y*b in [1/2, 1)), the value
// q + r*y lies strictly between q and the true quotient whenever they differ,
// hence rounds exactly as the true quotient does; r rounded toward the target
// keeps the sign needed when the remainder is not representable.
#define CPL_DIV_FIX(NAME, MODE)                                                              \
  __device__ __forceinline__ double NAME(double q, double b, double a, double y) {          \
    double r, res;                                                                          \
    asm("s_nop 1\n\t" MODE "v_fma_f64 %1, -%2, %3, %4\n\t"                                   \
        "v_fma_f64 %0, %1, %5, %2\n\t" CPL_MODE_RN                                           \
        : "=&v"(res), "=&v"(r)                                                              \
        : "v"(q), "v"(b), "v"(a), "v"(y));                                                  \
    return res;                                                                             \
  }
CPL_DIV_FIX(div_fix_rd, CPL_MODE_RD)
CPL_DIV_FIX(div_fix_ru, CPL_MODE_RU)
#undef CPL_DIV_FIX

// Scalar multiply of a possibly-tiny double by a power of two in the target
// mode (used to bring a scaled quotient back into the subnormal range).
#define CPL_SCALE_DIR(NAME, MODE)                                                            \
  __device__ __forceinline__ double NAME(double x, double p2) {                              \
    double r;                                                                               \
    asm("s_nop 1\n\t" MODE "v_mul_f64 %0, %1, %2\n\ts_nop 0\n\t" CPL_MODE_RN                 \
        : "=v"(r)                                                                           \
        : "v"(x), "v"(p2));                                                                 \
    return r;                                                                               \
  }
CPL_SCALE_DIR(scale_rd, CPL_MODE_RD)
CPL_SCALE_DIR(scale_ru, CPL_MODE_RU)
#undef CPL_SCALE_DIR

// Slow path: zeros, infinities, NaN, overflow, subnormal operands or results.
// Both operands are scaled by exact powers of two into [1, 2), the scaled
// quotient is rounded in the target mode (exact remainder there), and the
// power of two is re-applied in the target mode.  Rounding twice in the same
// direction onto nested grids equals rounding once, so underflowing and
// overflowing quotients come out exactly as IEEE directed division gives them.
template <bool UP>
__device__ __forceinline__ double div_slow(double a, double b) {
  double q = a / b;
  if (__builtin_isnan(q) || a == 0.0 || __builtin_isinf(a) || b == 0.0 || __builtin_isinf(b)) return q;
  if (b < 0.0) { a = -a; b = -b; }
  int shift = 0;  // true quotient = (a / b) * 2^shift after the scalings below
  if (dexp_field(b) == 0) { b *= 0x1p+600; shift += 600; }
  if (dexp_field(a) == 0) { a *= 0x1p+600; shift -= 600; }
  int eb = dexp_field(b) - 1023, ea = dexp_field(a) - 1023;
  b = __builtin_ldexp(b, -eb);  // exact: the results are normal
  a = __builtin_ldexp(a, -ea);
  shift += ea - eb;
  double qs = a / b;  // |qs| in (1/2, 2)
  double ds = UP ? div_fix_ru(qs, b, a, 0.5) : div_fix_rd(qs, b, a, 0.5);
  while (shift > 1000) { ds = UP ? scale_ru(ds, 0x1p+1000) : scale_rd(ds, 0x1p+1000); shift -= 1000; }
  while (shift < -1000) { ds = UP ? scale_ru(ds, 0x1p-1000) : scale_rd(ds, 0x1p-1000); shift += 1000; }
  return UP ? scale_ru(ds, pow2(shift)) : scale_rd(ds, pow2(shift));
}

// Fast-path condition: |q| and |b| in [2^-480, 2^480), q faithful.  Then
// |a| = |q b|(1+e) with |e| < 2^-51 lies in [2^-961, 2^961], so the remainder
// a - q*b of the faithful q is exactly representable, and y = +-2^-600 (the
// sign of b) satisfies 0 < y*b <= 2^-120 < 1: q + r*y is strictly between q
// and the true quotient (the fma computes r*y exactly, however small), so it
// rounds in the target direction to q when the quotient lies on the target
// side of q and to q's neighbour on that side otherwise.
__device__ __forceinline__ bool div_fast_ok(double q, double b) {
  double aq = __builtin_fabs(q), ab = __builtin_fabs(b);
  return (aq >= 0x1p-480) & (aq < 0x1p+480) & (ab >= 0x1p-480) & (ab < 0x1p+480);
}
__device__ __forceinline__ double div_y(double b) { return __builtin_copysign(0x1p-600, b); }

// Quotient for the fast path: the hardware reciprocal estimate, one Newton
// step (r <- r + r (1 - b r)) and one residual correction
// q1 = q0 + (a - b q0) r, all in round-to-nearest.  The correction squares the
// reciprocal's relative error e: q1 is faithful (one of the two doubles around
// a/b) whenever e^2 stays below about 2^-54 and q1 and b pass div_fast_ok, and
// the fix step below needs no more, since the remainder of a faithful quotient
// is exactly representable.  The estimate is accurate to 2^-24.2 (measured
// over every 2^-24-spaced mantissa at seven exponents), so e <= 2^-48 after the
// Newton step; one step suffices for any estimate better than 2^-13.5.  Outside
// the fast range the result is not used.
__device__ __forceinline__ double div_faithful(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  const double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q0 = a * r;
  const double rem = __builtin_fma(-b, q0, a);
  return __builtin_fma(rem, r, q0);
}

// Quotients the fast path does not take: 0/b is the exact signed zero for a
// finite nonzero b (IEEE division otherwise: NaN for 0/0, signed zero for
// 0/inf); every other operand pair goes through the scaled slow path.
template <bool UP>
__device__ __forceinline__ double div_rare(double a, double b) {
  if (a == 0.0) return (__builtin_isfinite(b) && b != 0.0) ? a * b : a / b;
  return div_slow<UP>(a, b);
}

template <bool UP>
__device__ __forceinline__ double div_dir(double a, double b) {
  double q = div_faithful(a, b);
  if (__builtin_expect(div_fast_ok(q, b), 1)) return UP ? div_fix_ru(q, b, a, div_y(b)) : div_fix_rd(q, b, a, div_y(b));
  return div_rare<UP>(a, b);
}
__device__ __forceinline__ double div_rd(double a, double b) { return div_dir<false>(a, b); }
__device__ __forceinline__ double div_ru(double a, double b) { return div_dir<true>(a, b); }

// The correction statements of div_pair2: both remainders, then both fixes,
// in one mode (SAME) or in two modes (M1 for the first, M2 for the second).
#define CPL_DIV2_SAME(M)                                                                     \
  asm("s_nop 1\n\t" M "v_fma_f64 %2, -%4, %5, %6\n\t"                                      \
      "v_fma_f64 %3, -%8, %9, %10\n\tv_fma_f64 %0, %2, %7, %4\n\t"                          \
      "v_fma_f64 %1, %3, %11, %8\n\t" CPL_MODE_RN                                            \
      : "=&v"(q1), "=&v"(q2), "=&v"(r1), "=&v"(r2)                                          \
      : "v"(f1), "v"(b1), "v"(a1), "v"(y1), "v"(f2), "v"(b2), "v"(a2), "v"(y2))
#define CPL_DIV2_TWO(M1, M2)                                                                 \
  asm("s_nop 1\n\t" M1 "v_fma_f64 %2, -%4, %5, %6\n\t"                                     \
      "v_fma_f64 %0, %2, %7, %4\n\t" M2 "v_fma_f64 %3, -%8, %9, %10\n\t"                   \
      "v_fma_f64 %1, %3, %11, %8\n\t" CPL_MODE_RN                                            \
      : "=&v"(q1), "=&v"(q2), "=&v"(r1), "=&v"(r2)                                          \
      : "v"(f1), "v"(b1), "v"(a1), "v"(y1), "v"(f2), "v"(b2), "v"(a2), "v"(y2))
// Two independent directed quotients a1 / b1 (upward if UP1) and a2 / b2
// (upward if UP2): the faithful quotients interleave, and both corrections run
// in one statement (two mode switches fewer than two separate divisions).  The
// same operations per quotient as div_dir, so the same bits.
template <bool UP1, bool UP2>
__device__ __forceinline__ void div_pair2(double a1, double b1, double a2, double b2, double& q1, double& q2) {
  const double f1 = div_faithful(a1, b1), f2 = div_faithful(a2, b2);
  if (__builtin_expect(div_fast_ok(f1, b1) & div_fast_ok(f2, b2), 1)) {
    const double y1 = div_y(b1), y2 = div_y(b2);
    double r1, r2;
    if constexpr (UP1 && UP2) CPL_DIV2_SAME(CPL_MODE_RU);
    else if constexpr (!UP1 && !UP2) CPL_DIV2_SAME(CPL_MODE_RD);
    else if constexpr (UP1) CPL_DIV2_TWO(CPL_MODE_RU, CPL_MODE_RD);
    else CPL_DIV2_TWO(CPL_MODE_RD, CPL_MODE_RU);
    return;
  }
  q1 = div_fast_ok(f1, b1) ? (UP1 ? div_fix_ru(f1, b1, a1, div_y(b1)) : div_fix_rd(f1, b1, a1, div_y(b1)))
                           : div_rare<UP1>(a1, b1);
  q2 = div_fast_ok(f2, b2) ? (UP2 ? div_fix_ru(f2, b2, a2, div_y(b2)) : div_fix_rd(f2, b2, a2, div_y(b2)))
                           : div_rare<UP2>(a2, b2);
}
#undef CPL_DIV2_SAME
#undef CPL_DIV2_TWO

// [rd(al / bl), ru(au / bu)]: both corrections in one statement.
__device__ __forceinline__ void div_pair(double al, double bl, double au, double bu, double& l, double& u) {
  div_pair2<false, true>(al, bl, au, bu, l, u);
}

// ---- square root ------------------------------------------------------------
// s = RN(sqrt x); r = fma(-s, s, x) and fma(r, y, s) in the target mode with
// y = 2^-(e_s+2), so r*y lies between 0 and sqrt(x) - s.
#define CPL_SQRT_FIX(NAME, MODE)                                                             \
  __device__ __forceinline__ double NAME(double s, double x, double y) {                    \
    double r, res;                                                                          \
    asm("s_nop 1\n\t" MODE "v_fma_f64 %1, -%2, %2, %3\n\t"                                   \
        "v_fma_f64 %0, %1, %4, %2\n\t" CPL_MODE_RN                                           \
        : "=&v"(res), "=&v"(r)                                                              \
        : "v"(s), "v"(x), "v"(y));                                                          \
    return res;                                                                             \
  }
CPL_SQRT_FIX(sqrt_fix_rd, CPL_MODE_RD)
CPL_SQRT_FIX(sqrt_fix_ru, CPL_MODE_RU)
#undef CPL_SQRT_FIX

template <bool UP>
__device__ __forceinline__ double sqrt_dir(double x) {
  if (!(x > 0.0) || __builtin_isinf(x)) return __builtin_sqrt(x);  // 0, -0, inf, NaN, negative
  int k = 0;
  if (dexp_field(x) < 64) { x *= 0x1p+200; k = -100; }  // tiny or subnormal: sqrt(x 2^200) = sqrt(x) 2^100
  double s = __builtin_sqrt(x);
  double y = pow2(1023 - 2 - dexp_field(s));
  double r = UP ? sqrt_fix_ru(s, x, y) : sqrt_fix_rd(s, x, y);
  if (k) r = UP ? scale_ru(r, pow2(k)) : scale_rd(r, pow2(k));
  return r;
}
__device__ __forceinline__ double sqrt_rd(double x) { return sqrt_dir<false>(x); }
__device__ __forceinline__ double sqrt_ru(double x) { return sqrt_dir<true>(x); }

}  // namespace cpl

// ============================================================================
// FP32 primitives for the outward-rounded prefilter.  MODE[1:0] is the FP32
// rounding field.  The prefilter only needs enclosures (its results are never
// reported), so operands outside the comfortable range widen to +-inf instead
// of taking an exact slow path.
// ============================================================================
namespace cpl {

#define CPL_MODE32_RD "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 2\n\t"
#define CPL_MODE32_RU "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 1\n\t"
#define CPL_MODE32_RN "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 0\n\t"

#define CPL_DIR_BINOP32(NAME, MODE, INSN)                                                    \
  __device__ __forceinline__ float NAME(float a, float b) {                                 \
    float r;                                                                                \
    asm("s_nop 1\n\t" MODE INSN " %0, %1, %2\n\ts_nop 0\n\t" CPL_MODE32_RN                  \
        : "=v"(r)                                                                           \
        : "v"(a), "v"(b));                                                                  \
    return r;                                                                               \
  }
CPL_DIR_BINOP32(add_rd, CPL_MODE32_RD, "v_add_f32")
CPL_DIR_BINOP32(add_ru, CPL_MODE32_RU, "v_add_f32")
CPL_DIR_BINOP32(mul_rd, CPL_MODE32_RD, "v_mul_f32")
CPL_DIR_BINOP32(mul_ru, CPL_MODE32_RU, "v_mul_f32")
#undef CPL_DIR_BINOP32

__device__ __forceinline__ void add_pair(float al, float bl, float au, float bu, float& l, float& u) {
  asm("s_nop 1\n\t" CPL_MODE32_RD "v_add_f32 %0, %2, %3\n\ts_nop 0\n\t" CPL_MODE32_RU
      "v_add_f32 %1, %4, %5\n\ts_nop 0\n\t" CPL_MODE32_RN
      : "=&v"(l), "=v"(u)
      : "v"(al), "v"(bl), "v"(au), "v"(bu));
}
__device__ __forceinline__ void sub_pair(float al, float bu, float au, float bl, float& l, float& u) {
  asm("s_nop 1\n\t" CPL_MODE32_RD "v_sub_f32 %0, %2, %3\n\ts_nop 0\n\t" CPL_MODE32_RU
      "v_sub_f32 %1, %4, %5\n\ts_nop 0\n\t" CPL_MODE32_RN
      : "=&v"(l), "=v"(u)
      : "v"(al), "v"(bu), "v"(au), "v"(bl));
}
__device__ __forceinline__ void mul_pair(float al, float bl, float au, float bu, float& l, float& u) {
  asm("s_nop 1\n\t" CPL_MODE32_RD "v_mul_f32 %0, %2, %3\n\ts_nop 0\n\t" CPL_MODE32_RU
      "v_mul_f32 %1, %4, %5\n\ts_nop 0\n\t" CPL_MODE32_RN
      : "=&v"(l), "=v"(u)
      : "v"(al), "v"(bl), "v"(au), "v"(bu));
}

// Division: q = RN(a/b) (IEEE), then fma(-q, b, a) and fma(r, y, q) in the
// target mode with y = +-2^-80: the result is on the correct side of the true
// quotient even when q is only faithful.  Outside |q|, |b| in [2^-50, 2^50]
// the bound widens to -inf / +inf.
__device__ __forceinline__ bool div_fast_ok32(float q, float b) {
  float aq = __builtin_fabsf(q), ab = __builtin_fabsf(b);
  return (aq >= 0x1p-50f) & (aq < 0x1p+50f) & (ab >= 0x1p-50f) & (ab < 0x1p+50f);
}
__device__ __forceinline__ void div_pair(float al, float bl, float au, float bu, float& l, float& u) {
  float ql = al / bl, qu = au / bu;
  float yl = __builtin_copysignf(0x1p-80f, bl), yu = __builtin_copysignf(0x1p-80f, bu), rl, ru;
  asm("s_nop 1\n\t" CPL_MODE32_RD "v_fma_f32 %2, -%4, %5, %6\n\t"
      "v_fma_f32 %0, %2, %7, %4\n\t" CPL_MODE32_RU "v_fma_f32 %3, -%8, %9, %10\n\t"
      "v_fma_f32 %1, %3, %11, %8\n\t" CPL_MODE32_RN
      : "=&v"(l), "=&v"(u), "=&v"(rl), "=&v"(ru)
      : "v"(ql), "v"(bl), "v"(al), "v"(yl), "v"(qu), "v"(bu), "v"(au), "v"(yu));
  if (!div_fast_ok32(ql, bl)) l = (al == 0.0f && bl != 0.0f) ? ql : -__builtin_huge_valf();
  if (!div_fast_ok32(qu, bu)) u = (au == 0.0f && bu != 0.0f) ? qu : __builtin_huge_valf();
}
__device__ __forceinline__ float div_rd(float a, float b) {
  float l, u;
  div_pair(a, b, a, b, l, u);
  return l;
}
__device__ __forceinline__ float div_ru(float a, float b) {
  float l, u;
  div_pair(a, b, a, b, l, u);
  return u;
}

// Outward conversion of a double to float.
__device__ __forceinline__ float next_down_f(float f) {
  if (f != f || f == -__builtin_huge_valf()) return f;
  if (f == 0.0f) return -0x1p-149f;
  return __int_as_float(__float_as_int(f) + (f > 0.0f ? -1 : 1));
}
__device__ __forceinline__ float f32_down(double x) {
  float f = (float)x;  // round to nearest
  return ((double)f > x) ? next_down_f(f) : f;
}
__device__ __forceinline__ float f32_up(double x) { return -f32_down(-x); }

}  // namespace cpl
