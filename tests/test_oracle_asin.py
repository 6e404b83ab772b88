"""The prover's own arcsine: accuracy against mpmath, enclosure of the widened
bounds, and identity of the coefficient table across generator, oracle and
HIP source."""
import os
import re

import mpmath as mp
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _points():
    rng = np.random.default_rng(2)
    return np.concatenate([rng.uniform(0, 1, 6000), rng.uniform(0.495, 0.505, 1500), 1 - rng.uniform(0, 1e-4, 1500),
                           rng.uniform(0, 1e-6, 300), np.linspace(0, 1, 1001), [0.5, np.nextafter(0.5, 1), 1.0]])


def test_asin_error_below_1_6_ulp(cplo):
    """1.50 ulp worst measured (near x = 0.506: the square root's own rounding error
    is not corrected); the reference's five-ulp widening covers any error below
    (5 - 1) / 2 = 2 ulps, with a step across a binade edge included."""
    mp.mp.dps = 40
    x = _points()
    y = cplo.prim("asin_rn", x)
    worst = 0.0
    for xi, yi in zip(x, y):
        t = mp.asin(mp.mpf(float(xi)))
        ulp = np.spacing(abs(float(t))) if t != 0 else 5e-324
        worst = max(worst, float(abs(mp.mpf(float(yi)) - t) / ulp))
    assert worst < 1.6, worst


def test_widened_bounds_enclose_true_asin(cplo):
    mp.mp.dps = 40
    x = np.concatenate([_points(), -_points()[:3000]])
    lo, hi = cplo.prim("dasin_rd", x), cplo.prim("dasin_ru", x)
    for xi, l, h in zip(x, lo, hi):
        t = mp.asin(mp.mpf(float(xi)))
        assert mp.mpf(float(l)) <= t <= mp.mpf(float(h)), xi
    # clamps of interval.cuh:426-437
    assert np.all(hi <= 1.5707963267948968) and np.all(lo >= -1.5707963267948968)
    assert np.all(lo[x >= 0] >= 0) and np.all(hi[x <= 0] <= 0)


def test_asin_symmetry_and_specials(cplo):
    x = _points()
    assert np.array_equal(cplo.prim("asin_rn", -x), -cplo.prim("asin_rn", x))
    s = cplo.prim("asin_rn", np.array([1.0, 0.0, 1.5, -2.0]))
    assert s[0] == 1.5707963267948966 and s[1] == 0.0 and np.isnan(s[2]) and np.isnan(s[3])


def _hexfloats(text):
    return [float.fromhex(h) for h in re.findall(r"-?0x1\.[0-9a-f]+p[-+]?\d+", text)]


def test_coefficient_tables_identical():
    from circlepacking_amd.csrc.gen_asin_poly import coefficients
    gen, _ = coefficients()
    oracle_src = open(os.path.join(ROOT, "oracle", "cpl_oracle.hpp")).read()
    hip_src = open(os.path.join(ROOT, "circlepacking_amd", "csrc", "cpl_asin.hpp")).read()
    o = oracle_src[oracle_src.index("asin_poly[13]"):oracle_src.index("asin_pio2_hi")]
    h = hip_src[hip_src.index("asin_poly_eval"):hip_src.index("return p;")]
    assert _hexfloats(o) == gen
    assert _hexfloats(h) == gen
