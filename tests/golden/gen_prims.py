"""Writes tests/golden/prims.npz: directed-rounding golden vectors.

Every expected result is computed from exact rational arithmetic
(fractions.Fraction) and rounded toward -inf / +inf by exact comparison, so
the vectors pin IEEE-754 directed add/sub/mul/div/sqrt independently of any
machine's rounding modes.  Run from the repository root:
    python tests/golden/gen_prims.py
"""
import math
import os
import sys
from fractions import Fraction

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from tests.operands import binary_operands  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "prims.npz")
DBL_MAX = Fraction(np.finfo(np.float64).max)


def round_dir(q, up):
    """Round the rational q to a double toward +inf (up) or -inf."""
    if q > DBL_MAX:
        return math.inf if up else float(DBL_MAX)
    if q < -DBL_MAX:
        return -float(DBL_MAX) if up else -math.inf
    f = float(q)  # correctly rounded to nearest (Python int/int division)
    fq = Fraction(f)
    if fq == q:
        return f
    if up and fq < q:
        f = math.nextafter(f, math.inf)
    elif not up and fq > q:
        f = math.nextafter(f, -math.inf)
    return f


def signed_zero(kind, a, b, up):
    """IEEE sign of an exact zero result."""
    if kind == "add":
        if a == 0 and b == 0 and math.copysign(1, a) == math.copysign(1, b):
            return a
        return 0.0 if up else -0.0
    if kind in ("mul", "div"):
        neg = (math.copysign(1, a) < 0) != (math.copysign(1, b) < 0)
        return -0.0 if neg else 0.0
    return a


def directed(kind, a, b, up):
    if not (math.isfinite(a) and math.isfinite(b)):
        return None
    if kind == "div" and b == 0:
        return None
    fa, fb = Fraction(a), Fraction(b)
    if kind == "add":
        q = fa + fb
    elif kind == "sub":
        q = fa - fb
        if q == 0:
            return signed_zero("add", a, -b, up)
        return round_dir(q, up)
    elif kind == "mul":
        q = fa * fb
    else:
        q = fa / fb
    if q == 0:
        return signed_zero(kind, a, b, up)
    return round_dir(q, up)


def directed_sqrt(x, up):
    if not math.isfinite(x) or x < 0:
        return None
    if x == 0:
        return x
    s = math.sqrt(x)  # correctly rounded to nearest
    fx = Fraction(x)
    if Fraction(s) ** 2 == fx:
        return s
    if up and Fraction(s) ** 2 < fx:
        s = math.nextafter(s, math.inf)
    elif not up and Fraction(s) ** 2 > fx:
        s = math.nextafter(s, -math.inf)
    return s


def main():
    a, b = binary_operands(seed=20261203, n=3000)
    keep = np.isfinite(a) & np.isfinite(b) & (b != 0)
    a, b = a[keep], b[keep]
    out = {"a": a, "b": b}
    for kind in ("add", "sub", "mul", "div"):
        for up in (False, True):
            out[f"{kind}_{'ru' if up else 'rd'}"] = np.array([directed(kind, x, y, up) for x, y in zip(a, b)])
    x = np.abs(a)
    out["sqrt_x"] = x
    for up in (False, True):
        out[f"sqrt_{'ru' if up else 'rd'}"] = np.array([directed_sqrt(v, up) for v in x])
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
