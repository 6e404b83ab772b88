"""Generate the polynomials of the prover's own round-to-nearest sin, cos.

Method (own design: Cody-Waite reduction by pi/2 in three parts, kernels on
[-pi/4, pi/4] with coefficients by Chebyshev interpolation in mpmath):

  sin(r) = r + r z S(z),            z = r^2,  S(z) ~ (sin(sqrt z) - sqrt z) / (z sqrt z)
  cos(r) = 1 - z/2 + z^2 C(z),                C(z) ~ (cos(sqrt z) - 1 + z/2) / z^2

on z in [0, (pi/4)^2].  pi/2 = P1 + P2 + P3 with P1 = RN(pi/2), P2 = RN(pi/2 - P1),
P3 = RN(pi/2 - P1 - P2).

Run:  python gen_trig_poly.py  -> coefficient tables (highest degree first) as
C99 hex-float literals for cpl_trig.hpp.
"""
import mpmath as mp

DEG_S = 7
DEG_C = 7


def fs(z):
    if z == 0:
        return -mp.mpf(1) / 6
    s = mp.sqrt(z)
    return (mp.sin(s) - s) / (z * s)


def fc(z):
    if z == 0:
        return mp.mpf(1) / 24
    s = mp.sqrt(z)
    return (mp.cos(s) - 1 + z / 2) / (z * z)


def coefficients():
    mp.mp.dps = 60
    hi = (mp.pi / 4) ** 2
    s, es = mp.chebyfit(fs, [0, hi], DEG_S + 1, error=True)
    c, ec = mp.chebyfit(fc, [0, hi], DEG_C + 1, error=True)
    p1 = float(mp.pi / 2)
    p2 = float(mp.pi / 2 - mp.mpf(p1))
    p3 = float(mp.pi / 2 - mp.mpf(p1) - mp.mpf(p2))
    return [float(v) for v in s], float(es), [float(v) for v in c], float(ec), (p1, p2, p3)


def main():
    s, es, c, ec, p = coefficients()
    print(f"// sin kernel S, degree {DEG_S}, highest first; chebyfit abs error {es:.3e}")
    for v in s:
        print(f"    {v.hex()},  // {v!r}")
    print(f"// cos kernel C, degree {DEG_C}, highest first; chebyfit abs error {ec:.3e}")
    for v in c:
        print(f"    {v.hex()},  // {v!r}")
    print("// pi/2 = " + " + ".join(v.hex() for v in p))
    mp.mp.dps = 40
    print(f"// 2/pi = {float(2 / mp.pi).hex()}")


if __name__ == "__main__":
    main()
