"""The oracle's criteria against the real-valued area inequalities they bound,
evaluated in mpmath at points (tests/point_model.py): a PROVEN cell satisfies
its family's inequality at every sampled interior point of the admissible
domain; at point cells, the two sides of every area comparison enclose the
real disk area and c x used area and agree with them to rounding, and the
verdict is PROVEN exactly when the inequality holds."""
import numpy as np
import pytest

from circlepacking_amd import abi
from tests import criteria_checks as cc
from tests import exact_ref as X
from tests.cells import c1_grid_soa, grid_cells, random_kernel_body_cells

FAMILIES = list(range(6))


@pytest.fixture(scope="module")
def ref(cplo):
    return X.Reference(lambda x: float(cplo.prim("asin_rn", np.array([x]))[0]))


@pytest.fixture(scope="module")
def lbrsp(cplo):
    return lambda la: float(cplo.prim("lbrsp", np.array([la]))[0])


def soundness_cells(cplo, ref, family):
    """PROVEN-rich cell sets: random boxes, coarse reference-grid cells and,
    for the 2-disk families, the 2-parameter 256^2 grid at lambda = 0.5."""
    pool = random_kernel_body_cells(700 + family, 20000, family)
    proven = pool[(cplo.eval_kernel_body(family, pool) & 7) == abi.V_PROVEN][:250]
    sets = [proven]
    for k, sd in enumerate([(16, 16, 16, 8), (64, 32, 32, 8), (128, 64, 64, 16), abi.REF_SUBDIV[family]]):
        sets.append(grid_cells(ref, family, 300, 300 + 7 * family + k, sd))
    if family in (abi.START_END_2D, abi.END_2D):
        soa = c1_grid_soa(256)[:, ::97]
        cells = np.zeros((soa.shape[1], 8))
        cells[:, 0:6] = soa[0:6].T
        sets.append(cells)
    return np.concatenate(sets)


@pytest.mark.parametrize("family", FAMILIES)
def test_proven_cells_satisfy_the_inequality_at_interior_points(cplo, ref, lbrsp, family):
    cells = soundness_cells(cplo, ref, family)
    bad, checked = cc.soundness_violations(cplo.eval_kernel_body, family, cells, seed=family, lbrsp=lbrsp)
    assert not bad, bad[:3]
    assert checked >= 800, checked


@pytest.mark.parametrize("family", FAMILIES)
def test_point_cell_area_terms_match_the_real_values(cplo, lbrsp, family):
    points = np.concatenate([cc.domain_points(family, 1500, 1000 + family, lbrsp),
                             cc.wide_points(family, 1500, 2000 + family)])
    bad, compared = cc.term_mismatches(cplo.area_terms, family, points)
    assert not bad, bad[:3]
    assert compared >= 1500, compared


@pytest.mark.parametrize("family", FAMILIES)
def test_point_cells_proven_where_the_inequality_holds(cplo, lbrsp, family):
    """In the admissible domain the inequalities hold at every point (the
    reference's proof reduces to that); every point cell is PROVEN."""
    points = cc.domain_points(family, 1500, 3000 + family, lbrsp)
    bad, proven, critical = cc.point_disagreements(cplo.eval_kernel_body, family, points)
    assert not bad, bad[:3]
    assert proven >= 1400 and critical == 0, (proven, critical)
