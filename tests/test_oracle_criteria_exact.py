"""The oracle's kernel bodies, admissible ranges, synthetic clipping and
reference-grid search against an independent exact-rational restatement
written from the reference's sources (tests/exact_ref.py): verdict bytes,
range endpoints, first problematic la, failure codes and printed critical-cell
lines must be identical.  The restatement shares only the round-to-nearest
arcsine with the oracle."""
import numpy as np
import pytest

from circlepacking_amd import abi
from tests import exact_ref as X
from tests.cells import grid_cells, random_kernel_body_cells

FAMILIES = list(range(6))


@pytest.fixture(scope="module")
def ref(cplo):
    return X.Reference(lambda x: float(cplo.prim("asin_rn", np.array([x]))[0]))


def _body(ref, f, c):
    return ref.kernel_body(f, X.IV(c[0], c[1]), X.IV(c[2], c[3]), X.IV(c[4], c[5]), X.IV(c[6], c[7]))


def stratified(bytes_, per_kind, rng):
    """Indices of up to per_kind cells of every distinct verdict byte."""
    pick = []
    for v in np.unique(bytes_):
        idx = np.nonzero(bytes_ == v)[0]
        pick.extend(rng.choice(idx, min(per_kind, idx.size), replace=False).tolist())
    return np.array(sorted(pick))


@pytest.mark.parametrize("family", FAMILIES)
def test_kernel_body_on_random_boxes(cplo, ref, family):
    """Random boxes around the admissible band, including degenerate, wide and
    out-of-order ones: from 40,000 boxes, up to 120 of every verdict byte the
    oracle assigns (every verdict kind, both failing criteria, each trap code,
    the start family's no-fit flag) are re-evaluated exactly."""
    pool = random_kernel_body_cells(900 + family, 40000, family)
    pool_bytes = cplo.eval_kernel_body(family, pool)
    cells = pool[stratified(pool_bytes, 120, np.random.default_rng(family))]
    want = cplo.eval_kernel_body(family, cells)
    got = np.array([_body(ref, family, c) for c in cells], dtype=np.uint8)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(cells[i].tolist(), int(want[i]), int(got[i])) for i in bad[:5]]
    kinds = set((want & 7).tolist())
    assert {abi.V_PROVEN, abi.V_CRITICAL, abi.V_TRAP, abi.V_EMPTY} <= kinds
    if family in (abi.MIDDLE, abi.END_2D, abi.END_3D):
        assert abi.V_NOFIT in kinds
    if family in (abi.START, abi.MIDDLE, abi.END_2D):
        assert {b >> 4 for b in want.tolist() if b & 7 == abi.V_CRITICAL} == {1, 2}


def area_term_mismatches(area_terms, ref, family, cells):
    """Cells whose area comparisons (the two sides of every check of the
    output_criticals evaluation) differ from the exact restatement's, bit for bit."""
    pairs, counts = area_terms(family, cells)
    bad = []
    for c, got, k in zip(cells, pairs, counts):
        want = ref.area_terms(family, X.IV(c[0], c[1]), X.IV(c[2], c[3]), X.IV(c[4], c[5]), X.IV(c[6], c[7]))
        if want is None:
            if k != -1:
                bad.append((c.tolist(), int(k), None))
        elif k != len(want) or not np.array_equal(np.array(want, dtype=np.float64).reshape(-1, 2).view(np.uint64),
                                                 got[:k].view(np.uint64)):
            bad.append((c.tolist(), got[:max(k, 0)].tolist(), want))
    return bad


@pytest.mark.parametrize("family", FAMILIES)
def test_area_terms_bit_exact(cplo, ref, family):
    """Both sides of every area comparison (output_criticals evaluation, every
    criterion) on random boxes of every verdict kind and reference-grid cells."""
    pool = random_kernel_body_cells(900 + family, 20000, family)
    cells = np.concatenate([pool[stratified(cplo.eval_kernel_body(family, pool), 60, np.random.default_rng(3))],
                            grid_cells(ref, family, 150, 8, (16, 16, 16, 8)),
                            grid_cells(ref, family, 150, 9, abi.REF_SUBDIV[family])])
    bad = area_term_mismatches(cplo.area_terms, ref, family, cells)
    assert not bad, bad[:3]


@pytest.mark.parametrize("family", FAMILIES)
def test_kernel_body_on_reference_grid_cells(cplo, ref, family):
    """Cells of the reference's own grid rule at the reference's subdivision
    counts and at coarser ones (wide cells, where criteria fail)."""
    full = abi.REF_SUBDIV[family]
    subdivs = [full, (64, 32, 32, 8), (16, 16, 16, 8), (6, 6, 6, 3)]
    seen = set()
    for k, sd in enumerate(subdivs):
        cells = grid_cells(ref, family, 250, 17 * family + k, sd)
        want = cplo.eval_kernel_body(family, cells)
        got = np.array([_body(ref, family, c) for c in cells], dtype=np.uint8)
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (sd, [(cells[i].tolist(), int(want[i]), int(got[i])) for i in bad[:5]])
        seen |= set(want.tolist())
    kinds = {b & 7 for b in seen}
    assert abi.V_PROVEN in kinds and abi.V_CRITICAL in kinds


@pytest.mark.parametrize("family", FAMILIES)
def test_admissible_ranges(cplo, ref, family):
    """compute_r1_range, the family's r2 range and its r3 range, bit for bit."""
    rng = np.random.default_rng(40 + family)
    n = 400
    la0 = rng.uniform(0.5, 0.99, n)
    law = rng.choice([0.0, 1e-9, 1e-4, 1e-2], n)
    lane = 1.0 - la0
    r1 = lane * rng.uniform(0.1, 0.5, n)
    r1w = lane * rng.choice([0.0, 1e-6, 1e-2, 0.1], n)
    r2 = r1 * rng.uniform(0.3, 1.0, n)
    r2w = lane * rng.choice([0.0, 1e-6, 1e-2], n)
    recs = np.stack([la0, np.minimum(la0 + law, X.LAMBDA_MAX), r1, r1 + r1w, r2, r2 + r2w], 1)
    got = cplo.ranges(family, recs)
    for rec, row in zip(recs, got):
        la, ra, rb = X.IV(rec[0], rec[1]), X.IV(rec[2], rec[3]), X.IV(rec[4], rec[5])
        a = ref.compute_r1_range(la)
        b = ref.compute_r2_range(family, la, ra)
        c = ref.compute_r3_range(family, la, ra, rb) if family in X.THREE_D else X.IV(0.0, 0.0)
        want = [a.lb, a.ub, b.lb, b.ub, c.lb, c.ub]
        assert np.array_equal(np.array(want).view(np.uint64), row.view(np.uint64)), (rec.tolist(), want, row.tolist())


@pytest.mark.parametrize("family", [abi.START, abi.START_END_3D, abi.MIDDLE, abi.END_3D, abi.START_END_2D])
@pytest.mark.parametrize("clip", [True, False])
def test_synthetic_cells(cplo, ref, family, clip):
    """Cells of the 4-parameter workload (root cells of a 64^4 grid and their
    bisected descendants): radii clipped to the admissible ranges, r3 from its
    position in the r3 range."""
    rng = np.random.default_rng(90 + family + 10 * clip)
    cfg = abi.cfg_4param()
    n = 500
    soa = np.empty((8, n))
    for k in range(n):
        for a in range(4):
            lo, hi = cfg.box_lb[a], cfg.box_ub[a]
            iv = X.get_subinterval(X.IV(lo, hi), int(rng.integers(64)), 64)
            for _ in range(int(rng.integers(0, 4))):
                iv = X.get_subinterval(iv, int(rng.integers(2)), 2)
            soa[2 * a, k], soa[2 * a + 1, k] = iv.lb, iv.ub
    want = cplo.eval_cells(family, soa, clip)
    got = np.array([ref.synthetic(family, X.IV(*soa[0:2, k]), X.IV(*soa[2:4, k]), X.IV(*soa[4:6, k]),
                                  X.IV(*soa[6:8, k]), clip) for k in range(n)], dtype=np.uint8)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(soa[:, i].tolist(), int(want[i]), int(got[i])) for i in bad[:5]]
    if clip:
        assert abi.V_OOD in set((want & 7).tolist())


MODE_R_CASES = [((abi.LAMBDA_MIN, abi.LAMBDA_MAX), (4, 4, 4, 2)), ((0.5, 0.6), (3, 4, 5, 3)),
                ((0.55, 0.6), (4, 3, 3, 2)), ((0.7, 0.71), (2, 5, 4, 3)), ((0.9, 0.99), (3, 3, 3, 2))]


@pytest.mark.parametrize("family", FAMILIES)
def test_reference_grid_search(cplo, ref, family):
    """One launch of the reference kernel on truncated grids: first problematic
    la, cells evaluated, failure rule (rc, trap code) and no-fit flags."""
    for outer, sd in MODE_R_CASES:
        if family not in X.THREE_D:
            sd = (sd[0], sd[1] + 2, 3 * sd[2], 1)
        want = cplo.mode_r(family, outer, sd)
        got = X.mode_r(ref, family, X.IV(*outer), sd)
        for key in ("rc", "first", "evaluated", "trap_code", "nofit_flags"):
            assert got[key] == want[key], (outer, sd, key, got[key], want[key])


@pytest.mark.parametrize("family", FAMILIES)
def test_reference_grid_output_mode(cplo, ref, family):
    """The output_criticals launch (the SIGINT re-run): one line per failing
    criterion of every cell, the start family's no-fit lines, and traps of
    every cell."""
    lines = 0
    for outer, sd in MODE_R_CASES:
        if family not in X.THREE_D:
            sd = (sd[0], sd[1] + 2, 3 * sd[2], 1)
        want = cplo.mode_r(family, outer, sd, output_criticals=True)
        got = X.mode_r(ref, family, X.IV(*outer), sd, output_criticals=True)
        for key in ("rc", "first", "evaluated", "trap_code", "nofit_flags"):
            assert got[key] == want[key], (outer, sd, key, got[key], want[key])
        assert want["n_records"] == len(got["records"])
        assert [tuple(r) for r in want["records"].tolist()] == got["records"]
        lines += len(got["records"])
    assert lines > 0


def test_empty_verdict_only_where_the_reference_checks(cplo, ref):
    """The kernel bodies return EMPTY when a tightened r1 or r2 is empty, in every
    family; the reference checks this in the end 3-disk family only
    (end_configuration.cu:232-234) and evaluates the criteria otherwise.  The
    difference is never visible: every range's upper bound is its parent's
    (r2 up to r1_.ub, r3 up to r2_.ub), so the tightened intervals of the
    reference grids and of clipped synthetic cells are not empty.  Checked on
    the 4-parameter workload (16^4 roots, refined to depth 6) and on cells of
    the reference grids of every family, at the reference's subdivision counts
    and at coarse ones."""
    st, _ = cplo.refine_run(abi.cfg_4param(grid=(16, 16, 16, 16)), want_flags=False)
    fv = st.as_dict()["family_verdicts"]
    assert sum(fv[r][f][abi.V_EMPTY] for r in range(st.rounds) for f in range(6)) == 0
    assert sum(fv[r][f][abi.V_PROVEN] for r in range(st.rounds) for f in range(6)) > 0
    for family in FAMILIES:
        for k, sd in enumerate([abi.REF_SUBDIV[family], (6, 6, 6, 3)]):
            kinds = cplo.eval_kernel_body(family, grid_cells(ref, family, 2000, 70 + 7 * family + k, sd)) & 7
            assert abi.V_EMPTY not in set(kinds.tolist()), (family, sd)
