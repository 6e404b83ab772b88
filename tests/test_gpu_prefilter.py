"""FP32 outward-rounded prefilter: its primitives against x86 directed FP32
arithmetic, its arcsine against the mathematical arcsine, its geometry as an
enclosure of the FP64 geometry, and refinement runs with the prefilter giving
results identical to the plain FP64 path."""
import mpmath as mp
import numpy as np
import pytest

from circlepacking_amd import abi
from tests.cells import random_kernel_body_cells

pytestmark = pytest.mark.gpu

ASIN32_STEPS = 6  # outward steps of the FP32 arcsine bounds (cpl_asin.hpp)


def f32_values(rng, n, lo_exp=-40, hi_exp=40):
    m = rng.uniform(1.0, 2.0, n)
    e = rng.integers(lo_exp, hi_exp, n)
    s = rng.choice([-1.0, 1.0], n)
    return (s * np.ldexp(m, e)).astype(np.float32).astype(np.float64)


def f32_bits(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).view(np.uint32)


@pytest.mark.parametrize("op", ["add_rd32", "add_ru32", "sub_rd32", "sub_ru32", "mul_rd32", "mul_ru32",
                                "div_rd32", "div_ru32"])
def test_prim32_bit_exact_in_range(cpl, cplo, op):
    rng = np.random.default_rng(sum(map(ord, op)))
    a = f32_values(rng, 500_000, -20, 20)
    b = f32_values(rng, 500_000, -20, 20)
    # near-cancellation and exact cases
    a[:1000] = b[:1000]
    a[1000:2000] = -b[1000:2000]
    a[2000:3000] = np.nextafter(b[2000:3000].astype(np.float32), np.float32(np.inf)).astype(np.float64)
    a[3000:3100] = 0.0
    got, want = cpl.prim(op, a, b), cplo.prim(op, a, b)
    bad = np.nonzero(f32_bits(got) != f32_bits(want))[0]
    assert bad.size == 0, (op, a[bad[0]], b[bad[0]], got[bad[0]], want[bad[0]])


@pytest.mark.parametrize("op", ["div_rd32", "div_ru32"])
def test_div32_outside_fast_range_is_outward(cpl, cplo, op):
    rng = np.random.default_rng(3)
    a = f32_values(rng, 200_000, -120, 120)
    b = f32_values(rng, 200_000, -120, 120)
    got, want = cpl.prim(op, a, b), cplo.prim(op, a, b)
    if op == "div_rd32":
        assert np.all(got <= want)
    else:
        assert np.all(got >= want)
    fast = (np.abs(want) >= 2.0 ** -50) & (np.abs(want) < 2.0 ** 50) & (np.abs(b) >= 2.0 ** -50) & (np.abs(b) < 2.0 ** 50)
    assert np.array_equal(f32_bits(got[fast]), f32_bits(want[fast]))


def test_f32_outward_conversion(cpl, cplo):
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-2, 2, 200_000), f32_values(rng, 10_000),
                        [0.0, -0.0, 0.5, 0.9900000000000001, 1e-300, -1e-300]])
    for op in ("f32_down", "f32_up"):
        got, want = cpl.prim(op, x), cplo.prim(op, x)
        assert np.array_equal(f32_bits(got), f32_bits(want)), op
    assert np.all(cpl.prim("f32_down", x) <= x) and np.all(cpl.prim("f32_up", x) >= x)


def _asin_inputs():
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-1, 1, 2_000_000),
        np.ldexp(rng.uniform(1, 2, 200_000), rng.integers(-126, 0, 200_000)),
        1.0 - np.ldexp(rng.uniform(1, 2, 200_000), rng.integers(-24, -1, 200_000)),
        [0.0, 0.5, -0.5, 1.0, -1.0, 1e-30, 0.49999997, 0.50000006],
    ]).astype(np.float32).astype(np.float64)
    return np.clip(x, -1.0, 1.0)


def test_asin32_error_bound(cpl):
    x = _asin_inputs()
    got = cpl.prim("asin32_rn", x)
    t = np.arcsin(x)
    ulp = np.spacing(np.abs(t).astype(np.float32)).astype(np.float64)
    err = np.abs(got - t) / ulp
    worst = float(err.max())
    # the outward bounds step ASIN32_STEPS ulps; the rounding error must stay well inside
    assert worst <= 2.0, worst
    assert ASIN32_STEPS >= 2 * worst + 1


def test_dasin32_encloses_arcsine(cpl):
    x = _asin_inputs()
    lo, hi = cpl.prim("dasin32_rd", x), cpl.prim("dasin32_ru", x)
    t = np.arcsin(x)
    assert np.all(lo <= t) and np.all(t <= hi)
    # spot-check the tightest cases against 40-digit arcsine
    mp.mp.dps = 40
    close = np.argsort(np.minimum(t - lo, hi - t) / np.maximum(np.abs(t), 1e-300))[:200]
    for i in close:
        v = mp.asin(mp.mpf(float(x[i])))
        assert mp.mpf(float(lo[i])) <= v <= mp.mpf(float(hi[i])), float(x[i])


def _boxes(seed, n):
    rng = np.random.default_rng(seed)
    la0 = rng.uniform(0.5, 0.98, n)
    law = rng.choice([0.0, 1e-5, 1e-3], n)
    lane = 1 - la0 - law
    ra0 = lane * rng.uniform(0.25, 0.45, n)
    rb0 = ra0 * rng.uniform(0.6, 1.0, n)
    w = lane * rng.choice([0.0, 1e-6, 1e-3, 1e-2], n)
    return np.stack([la0, la0 + law, ra0, ra0 + w, rb0, rb0 + w * 0.5], 1)


@pytest.mark.parametrize("fn", ["inner_inner", "outer_inner", "inner_outer", "outer_outer",
                                "semicircle_outer", "semicircle_inner"])
def test_geometry_fp64_bit_exact_and_fp32_encloses(cpl, cplo, fn):
    recs = _boxes(sum(map(ord, fn)), 20_000)
    g64, t64 = cpl.debug_geom(fn, recs, fp32=False)
    o64, ot = cplo.geom(fn, recs)
    ok = ot == 0
    assert np.array_equal(t64, ot)
    assert np.array_equal(g64[ok].view(np.uint64), o64[ok].view(np.uint64))
    g32, t32 = cpl.debug_geom(fn, recs, fp32=True)
    both = ok & (t32 == 0) & (o64[:, 0] <= o64[:, 1])
    # the wider FP32 intervals hit the reference's consistency traps more often
    # (a trap only hands the cell to FP64); most records still compare
    assert both.sum() > 0.8 * len(recs)
    # the FP32 enclosure contains the FP64 one (both enclose the true range; FP32 is the wider)
    assert np.all(g32[both, 0] <= o64[both, 0]), fn
    assert np.all(g32[both, 1] >= o64[both, 1]), fn
    # ... and every point value inside the box
    rng = np.random.default_rng(2)
    pts = recs.copy()
    for k in range(3):
        v = rng.uniform(recs[:, 2 * k], recs[:, 2 * k + 1])
        pts[:, 2 * k] = v
        pts[:, 2 * k + 1] = v
    p64, pt = cplo.geom(fn, pts)
    sel = both & (pt == 0)
    assert np.all(g32[sel, 0] <= p64[sel, 0]) and np.all(g32[sel, 1] >= p64[sel, 1])


@pytest.mark.parametrize("family", range(6))
def test_fp32_never_proves_what_fp64_does_not(cpl, family):
    cells = random_kernel_body_cells(100 + family, 200_000, family)
    r32, v64 = cpl.debug_prefilter(family, cells)
    v = v64 & 7
    proven32 = r32 == 0
    assert not np.any(proven32 & (v != abi.V_PROVEN)), np.nonzero(proven32 & (v != abi.V_PROVEN))[0][:5]
    # FP32 decides a large share of the cells FP64 proves
    n64 = int(np.sum(v == abi.V_PROVEN))
    if n64 > 1000:
        assert proven32.sum() > 0.5 * n64


@pytest.mark.parametrize("opts", [0, abi.OPT_GRAPH | abi.OPT_HIER_ROOTS, abi.OPT_PERSISTENT | abi.OPT_HIER_ROOTS])
def test_prefilter_runs_identical_to_plain(cpl, opts):
    base = abi.cfg_4param(grid=(16, 16, 16, 16))
    plain_s, plain_f = cpl.refine_run(base)
    cfg = abi.with_options(base, opts | abi.OPT_FP32)
    for _ in range(2):
        s, f = cpl.refine_run(cfg)
        assert s.as_dict() == plain_s.as_dict()
        assert np.array_equal(f, plain_f)
        settled = s.fp32_settled_total()
        proven = sum(s.verdicts[d][abi.V_PROVEN] for d in range(s.rounds))
        assert 0 < settled <= proven


def test_prefilter_depth_mask(cpl):
    base = abi.cfg_4param(grid=(16, 16, 16, 16))
    plain_s, plain_f = cpl.refine_run(base)
    s, f = cpl.refine_run(abi.with_options(base, abi.OPT_FP32, prefilter_depths=0b1))
    assert s.as_dict() == plain_s.as_dict() and np.array_equal(f, plain_f)
    assert sum(s.fp32_settled[0]) > 0
    assert all(sum(s.fp32_settled[d]) == 0 for d in range(1, s.rounds))


def test_prefilter_full_grid_identical(cpl):
    base = abi.cfg_4param()
    plain = cpl.refine_run(base)
    for opts in (abi.OPT_GRAPH | abi.OPT_HIER_ROOTS, abi.OPT_PERSISTENT | abi.OPT_HIER_ROOTS):
        s, f = cpl.refine_run(abi.with_options(base, opts | abi.OPT_FP32))
        assert s.as_dict() == plain[0].as_dict()
        assert np.array_equal(f, plain[1])
