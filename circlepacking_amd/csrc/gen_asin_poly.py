"""Generate the polynomial used by the prover's own round-to-nearest asin.

Method (own design, standard argument reduction, coefficients by Chebyshev
interpolation in mpmath at 60 digits):

  f(z) = (asin(sqrt z) - sqrt z) / (z sqrt z),   z in [0, 1/4],  f(0) = 1/6

  |x| <= 1/2 :  asin(x) = x + (x z) f(z),           z = x^2
  |x| >  1/2 :  asin(|x|) = pi/2 - 2 asin(s),        s = sqrt(t), t = (1 - |x|)/2
                           = pi/2 - 2 (s + (s t) f(t))

f is approximated by the degree-12 polynomial P interpolating f at the Chebyshev
nodes of [0, 1/4] (mpmath.chebyfit); its max relative error on [0, 1/4] is
about 9e-17 relative to f, i.e. < 0.04 ulp of the final asin.

The FP32 prefilter's arcsine uses the same reduction with the degree-5
interpolant rounded to float (abs error 4.2e-9 relative to f, far below its
2-ulp FP32 rounding error) and pi/2 split as float hi + float lo.

Run:  python gen_asin_poly.py [--fp32]  -> prints the coefficient table (highest
degree first) as C99 hex-float literals, ready to paste into cpl_asin.hpp.
"""
import sys

import mpmath as mp

DEGREE = 12
DEGREE32 = 5


def f(z):
    if z == 0:
        return mp.mpf(1) / 6
    s = mp.sqrt(z)
    return (mp.asin(s) - s) / (z * s)


def coefficients(degree=DEGREE):
    mp.mp.dps = 60
    poly, err = mp.chebyfit(f, [0, mp.mpf(1) / 4], degree + 1, error=True)
    return [float(c) for c in poly], float(err)


def main_fp32():
    import numpy as np

    coeffs, err = coefficients(DEGREE32)
    print(f"// degree {DEGREE32}, highest first, rounded to float; chebyfit abs error {err:.3e}")
    for c in coeffs:
        print(f"    {float(np.float32(c)).hex()}f")
    hi = float(np.float32(float(mp.pi / 2)))
    lo = float(np.float32(float(mp.pi / 2 - mp.mpf(hi))))
    print(f"// pi/2 hi {hi.hex()}f  lo {lo.hex()}f")


def main():
    if "--fp32" in sys.argv:
        return main_fp32()
    coeffs, err = coefficients()
    print(f"// degree {DEGREE}, highest first; chebyfit abs error {err:.3e}")
    for c in coeffs:
        print(f"    {c.hex()},  // {c!r}")
    pio2_hi = float(mp.pi / 2)
    pio2_lo = float(mp.pi / 2 - mp.mpf(pio2_hi))
    print(f"// pi/2 hi {pio2_hi.hex()}  lo {pio2_lo.hex()}")


if __name__ == "__main__":
    main()
