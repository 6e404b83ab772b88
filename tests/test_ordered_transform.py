"""Atomic_ordered_double's map (atomic_float.hpp:56-66), by brute force: the
oracle's transform equals the restatement of the reference's bit expression,
round-trips every bit pattern (NaNs included), and orders every pair of
non-NaN doubles as the doubles are ordered, so an unsigned atomicMin over the
images is the minimum over the doubles."""
import numpy as np
import pytest

from tests import exact_ref as X


def patterns():
    """Every double whose low 48 bits are zero (all signs, exponents and top
    mantissa nibbles: 65,536 patterns), the specials, and random doubles."""
    top = (np.arange(1 << 16, dtype=np.uint64) << np.uint64(48))
    special = np.array([0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, -2.2250738585072014e-308, 1.0, -1.0,
                        1.7976931348623157e308, -1.7976931348623157e308, np.inf, -np.inf, 0.5, 0.9900000000000001])
    rng = np.random.default_rng(0)
    rnd = rng.integers(0, 1 << 63, 100000, dtype=np.uint64) | (rng.integers(0, 2, 100000, dtype=np.uint64) << np.uint64(63))
    return np.concatenate([top, special.view(np.uint64), rnd])


def test_transform_matches_the_reference_expression(cplo):
    b = patterns()
    got = cplo.prim("ordered", b.view(np.float64)).view(np.uint64)
    want = np.array([X.ordered_from_bits(int(v)) for v in b], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_round_trip_every_pattern(cplo):
    b = patterns()
    img = cplo.prim("ordered", b.view(np.float64))
    back = cplo.prim("ordered_inv", img).view(np.uint64)
    assert np.array_equal(back, b)
    assert all(X.bits_from_ordered(X.ordered_from_bits(int(v))) == int(v) for v in b[::97])


def test_order_of_every_adjacent_pair(cplo):
    b = patterns()
    d = b.view(np.float64)
    keep = ~np.isnan(d)
    d, b = d[keep], b[keep]
    # sort by value, with -0 before +0 (the map orders them so)
    order = np.lexsort((~(b >> np.uint64(63)).astype(bool), d))
    d, b = d[order], b[order]
    img = cplo.prim("ordered", b.view(np.float64)).view(np.uint64)
    same = b[1:] == b[:-1]
    assert np.all((img[1:] > img[:-1]) | same)
    assert np.all(np.diff(d)[~same] >= 0)


@pytest.mark.parametrize("seed", range(3))
def test_unsigned_minimum_is_the_double_minimum(cplo, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(0, 1, 5000) * 10.0 ** rng.integers(-300, 300, 5000)
    img = cplo.prim("ordered", v).view(np.uint64)
    m = cplo.prim("ordered_inv", np.array([img.min()]).view(np.float64))[0]
    assert m == v.min()
