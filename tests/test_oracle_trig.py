"""The rest of the interval surface in the oracle: our round-to-nearest sin,
cos, acos against 40-digit mpmath (error bound that the 3-ulp widening must
cover), and interval sin / cos / acos / sqrt / reciprocal / scalar division as
enclosures of the mathematical functions at interior points."""
import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 40


def ulp_err(got, exact):
    e = abs(mp.mpf(float(got)) - exact)
    u = mp.mpf(float(np.spacing(abs(float(exact))))) if exact != 0 else mp.mpf(2) ** -1074
    return float(e / u)


def near_multiples():
    pts = []
    for k in range(0, 6):
        c = float(mp.pi * k / 2)
        for d in range(-300, 301):
            v = c + d * np.spacing(max(c, 1e-300))
            if 0 <= v <= 8:
                pts.append(v)
    return np.array(pts)


@pytest.mark.parametrize("op", ["sin_rn", "cos_rn"])
def test_sin_cos_error_below_one_ulp(cplo, op):
    f = mp.sin if op == "sin_rn" else mp.cos
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 7, 200_000), np.linspace(0, 8, 4001), near_multiples()])
    got = cplo.prim(op, x)
    ref = np.sin(x) if op == "sin_rn" else np.cos(x)
    cand = np.argsort(-np.abs(got - ref) / np.spacing(np.abs(ref) + 1e-300))[:400]
    cand = np.union1d(cand, np.arange(len(x) - len(near_multiples()), len(x)))
    worst = max(ulp_err(got[i], f(mp.mpf(float(x[i])))) for i in cand)
    assert worst < 1.0, worst


def test_acos_error_below_one_ulp(cplo):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1, 1, 200_000),
                        1 - np.ldexp(rng.uniform(1, 2, 20_000), rng.integers(-53, -1, 20_000)),
                        -1 + np.ldexp(rng.uniform(1, 2, 20_000), rng.integers(-53, -1, 20_000)),
                        [-1.0, 1.0, 0.5, -0.5, 0.0, np.nextafter(0.5, 1), np.nextafter(-0.5, -1)]])
    got = cplo.prim("acos_rn", x)
    ref = np.arccos(x)
    cand = np.argsort(-np.abs(got - ref) / np.spacing(np.abs(ref) + 1e-300))[:400]
    worst = max(ulp_err(got[i], mp.acos(mp.mpf(float(x[i])))) for i in cand)
    assert worst < 1.0, worst
    assert got[x == 1.0][0] == 0.0


def _random_intervals(rng, n, lo, hi, widths):
    a = rng.uniform(lo, hi, n)
    w = rng.choice(widths, n) * rng.uniform(0, 1, n)
    return np.stack([a, a + w], 1)


@pytest.mark.parametrize("op,fn", [("sin", mp.sin), ("cos", mp.cos)])
def test_interval_sin_cos_enclose(cplo, op, fn):
    rng = np.random.default_rng(3)
    iv = np.concatenate([
        _random_intervals(rng, 3000, -10, 10, [0.0, 1e-6, 1e-2, 0.5, 3.0]),
        _random_intervals(rng, 500, -1e4, 1e4, [0.0, 1e-3, 2.0, 10.0]),
        np.array([[0.0, 7.0], [-0.1, 0.1], [1.5, 1.6], [3.1, 3.2], [4.7, 4.8], [6.2, 6.3], [-4.0, -3.0]]),
    ])
    out, trap = cplo.interval_op(op, iv)
    assert np.all(trap == 0)
    assert np.all(out[:, 0] <= out[:, 1]) and np.all(out[:, 0] >= -1) and np.all(out[:, 1] <= 1)
    for (a, b), (lo, hi) in zip(iv[::7], out[::7]):
        for p in np.linspace(a, b, 5):
            v = fn(mp.mpf(float(p)))
            assert mp.mpf(float(lo)) <= v <= mp.mpf(float(hi)), (op, a, b, p, lo, hi)
    # a full period gives [-1, 1]
    full, _ = cplo.interval_op(op, np.array([[0.0, 6.3], [-100.0, -90.0]]))
    assert np.array_equal(full, [[-1.0, 1.0], [-1.0, 1.0]])


def test_interval_acos_sqrt_reciprocal_divscalar_enclose(cplo):
    rng = np.random.default_rng(4)
    iv = _random_intervals(rng, 2000, -1, 0.99, [0.0, 1e-8, 1e-3, 0.5])
    iv[:, 1] = np.minimum(iv[:, 1], 1.0)
    out, trap = cplo.interval_op("acos", iv)
    assert np.all(trap == 0)
    for (a, b), (lo, hi) in zip(iv[::5], out[::5]):
        for p in (a, b, 0.5 * (a + b)):
            v = mp.acos(mp.mpf(float(p)))
            assert mp.mpf(float(lo)) <= v <= mp.mpf(float(hi))
    _, t = cplo.interval_op("acos", np.array([[-1.5, 0.0], [0.0, 1.0000000000000002]]))
    assert np.all(t == 1)
    pos = _random_intervals(rng, 2000, 1e-3, 1e3, [0.0, 1.0, 100.0])
    out, trap = cplo.interval_op("sqrt", pos)
    assert np.all(trap == 0)
    assert all(mp.mpf(float(lo)) ** 2 <= mp.mpf(float(a)) and mp.mpf(float(hi)) ** 2 >= mp.mpf(float(b))
               for (a, b), (lo, hi) in zip(pos, out))
    _, t = cplo.interval_op("sqrt", np.array([[-1e-300, 1.0]]))
    assert t[0] == 5
    neg = -pos[:, ::-1]
    for x in (pos, neg):
        out, _ = cplo.interval_op("reciprocal", x)
        assert np.all(out[:, 0] <= out[:, 1])
        assert all(mp.mpf(float(lo)) <= 1 / mp.mpf(float(p)) <= mp.mpf(float(hi))
                   for (a, b), (lo, hi) in zip(x[::10], out[::10]) for p in (a, b))
    out, _ = cplo.interval_op("reciprocal", np.array([[-1.0, 2.0]]))
    assert out[0, 0] == -np.inf and out[0, 1] == np.inf
    for s in (3.0, -7.0):
        out, _ = cplo.interval_op("div_scalar", pos, np.array([s, s]))
        assert all(mp.mpf(float(lo)) <= mp.mpf(float(p)) / s <= mp.mpf(float(hi))
                   for (a, b), (lo, hi) in zip(pos[::10], out[::10]) for p in (a, b))
