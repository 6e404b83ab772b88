"""The oracle's directed-rounding primitives against exact-rational golden vectors."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "prims.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("op", ["add_rd", "add_ru", "sub_rd", "sub_ru", "mul_rd", "mul_ru", "div_rd", "div_ru"])
def test_binary_ops_match_exact_rational_rounding(cplo, golden, op):
    a, b, want = golden["a"], golden["b"], golden[op]
    got = cplo.prim(op, a, b)
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, f"{op}: a={a[bad[0]]!r} b={b[bad[0]]!r} got {got[bad[0]]!r} want {want[bad[0]]!r}"


@pytest.mark.parametrize("op", ["sqrt_rd", "sqrt_ru"])
def test_sqrt_matches_exact_rational_rounding(cplo, golden, op):
    x, want = golden["sqrt_x"], golden[op]
    got = cplo.prim(op, x)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_directed_pairs_bracket_and_are_adjacent(cplo, golden):
    """rd <= ru, and they are equal or adjacent doubles (a planted RD/RU swap fails here)."""
    a, b = golden["a"], golden["b"]
    for op in ("add", "mul", "div"):
        lo, hi = cplo.prim(op + "_rd", a, b), cplo.prim(op + "_ru", a, b)
        ok = np.isfinite(lo) & np.isfinite(hi)
        assert np.all(lo[ok] <= hi[ok])
        gap = np.nextafter(lo[ok], np.inf)
        assert np.all((hi[ok] == lo[ok]) | (hi[ok] == gap))


def test_power_of_two_scalar_products_are_exact(cplo):
    """Interval x scalar rounds the upper bound down in the reference; at the path's
    factors 0.25, 0.5, 2 and normal operands both roundings give the same bits."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-4, 4, 100000), rng.uniform(1e-300, 1e-290, 1000)])
    for n in (0.25, 0.5, 2.0):
        assert np.array_equal(cplo.prim("mul_rd", x, n), cplo.prim("mul_ru", x, n))
