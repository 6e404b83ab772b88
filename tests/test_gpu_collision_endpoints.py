"""The collision-angle kernel specialised on the endpoint its caller reads
(compute_collision_angle_r<T, W>, collision_angles.cu:37-117): the lower-only
and upper-only forms give the bits of the full computation on the endpoint
they compute, with the same trap codes, and the full form matches the oracle.

Records are (lc, sc, ab) intervals built to cover every branch of the
reference function: definitely obtuse, definitely acute, obtuseness
undecided (both branches run; the first branch's result must then be computed
in full because its emptiness decides between taking and joining the second),
and the consistency traps (empty x or quotient intervals, arcsine domain).
Records are also evaluated grouped by class, so that whole waves take the
single-endpoint quotient paths that need every lane in the same case."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _records(seed, n):
    rng = np.random.default_rng(seed)
    lc0 = rng.uniform(0.3, 1.0, n)
    sc0 = lc0 * rng.uniform(0.2, 1.0, n)
    # ab around the right-angle value sqrt(lc^2 - sc^2): below it the angle at
    # the long side is obtuse, above it acute
    right = np.sqrt(np.maximum(lc0 * lc0 - sc0 * sc0, 1e-6))
    ab0 = right * rng.choice([0.5, 0.9, 0.99, 1.0, 1.01, 1.1, 1.6], n) * rng.uniform(0.97, 1.03, n)
    w = rng.choice([0.0, 1e-9, 1e-5, 1e-3, 3e-2, 0.2], (n, 3)) * rng.uniform(0.0, 1.0, (n, 3))
    recs = np.stack([lc0, lc0 + w[:, 0], sc0, sc0 + w[:, 1], ab0, ab0 + w[:, 2]], 1)
    # a few sc intervals longer than lc (tightening empties them) and tiny ab
    k = n // 50
    recs[:k, 2:4] = recs[:k, 0:2] + rng.uniform(0.0, 0.05, (k, 1))
    recs[k:2 * k, 4:6] = rng.uniform(1e-12, 1e-6, (k, 1)) * np.array([1.0, 1.5])
    return recs


def _obtuse_class(recs):
    """Approximate is_obtuse(lc, sc, ab) (collision_angles.cu:29-31) in plain
    floating point: 0 definitely acute, 1 undecided, 2 definitely obtuse."""
    lc_l, lc_u, sc_l, sc_u, ab_l, ab_u = recs.T
    lc_l = np.maximum(lc_l, sc_l)
    sc_u = np.minimum(sc_u, lc_u)
    definitely = lc_l ** 2 > sc_u ** 2 + ab_u ** 2
    possibly = lc_u ** 2 > sc_l ** 2 + ab_l ** 2
    return np.where(definitely, 2, np.where(possibly, 1, 0))


@pytest.mark.parametrize("grouped", [False, True])
def test_endpoint_forms_equal_full_computation_and_oracle(cpl, cplo, grouped):
    recs = _records(17, 60_000)
    cls = _obtuse_class(recs)
    if grouped:
        # waves of one obtuseness class: the wave-uniform single-endpoint
        # quotient paths (all lanes definitely obtuse / all acute) are taken
        order = np.argsort(cls, kind="stable")
        recs, cls = recs[order], cls[order]
    # every branch combination is represented
    for c in (0, 1, 2):
        assert np.count_nonzero(cls == c) > 2_000, c
    full, tf = cpl.debug_geom("compute", recs)
    lo, tl = cpl.debug_geom("compute_lb", recs)
    hi, th = cpl.debug_geom("compute_ub", recs)
    want, tw = cplo.geom("compute", recs)
    # identical trap decisions in all three forms and in the oracle
    assert np.array_equal(tf, tw)
    assert np.array_equal(tl, tw) and np.array_equal(th, tw)
    traps = np.unique(tw)
    assert len(traps) >= 3, traps  # no trap plus at least two trap kinds
    ok = tw == 0
    assert np.count_nonzero(ok & (cls == 1)) > 1_000
    bits = lambda a: np.ascontiguousarray(a).view(np.uint64)
    assert np.array_equal(bits(full[ok]), bits(want[ok]))
    assert np.array_equal(bits(lo[ok, 0]), bits(full[ok, 0]))
    assert np.array_equal(bits(hi[ok, 1]), bits(full[ok, 1]))


def test_endpoint_forms_fp32(cpl):
    recs = _records(29, 20_000)
    full, tf = cpl.debug_geom("compute", recs, fp32=True)
    lo, tl = cpl.debug_geom("compute_lb", recs, fp32=True)
    hi, th = cpl.debug_geom("compute_ub", recs, fp32=True)
    assert np.array_equal(tl, tf) and np.array_equal(th, tf)
    ok = tf == 0
    assert np.count_nonzero(ok) > 10_000
    bits = lambda a: np.ascontiguousarray(a).view(np.uint64)
    assert np.array_equal(bits(lo[ok, 0]), bits(full[ok, 0]))
    assert np.array_equal(bits(hi[ok, 1]), bits(full[ok, 1]))
