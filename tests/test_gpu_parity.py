"""Cell verdicts and refinement runs of the HIP path, identical to the CPU oracle."""
import os

import numpy as np
import pytest

from circlepacking_amd import abi
from tests.cells import c1_grid_soa, random_kernel_body_cells

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("family", range(6))
def test_kernel_body_random_cells(cpl, cplo, family):
    cells = random_kernel_body_cells(seed=100 + family, n=20000, family=family)
    got = cpl.eval_kernel_body(family, cells)
    want = cplo.eval_kernel_body(family, cells)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first cell {cells[bad[0]].tolist()} got {got[bad[0]]} want {want[bad[0]]}"
    # the test set exercises every verdict kind the family can produce
    kinds = set((want & 7).tolist())
    assert abi.V_PROVEN in kinds and abi.V_CRITICAL in kinds


@pytest.mark.parametrize("family", [abi.START_END_2D, abi.END_2D])
@pytest.mark.parametrize("clip", [True, False])
def test_small_grid_cells(cpl, cplo, family, clip):
    soa = c1_grid_soa(256)
    got = cpl.eval_cells(family, soa, clip)
    want = cplo.eval_cells(family, soa, clip)
    assert np.array_equal(got, want)
    golden = np.load(os.path.join(os.path.dirname(__file__), "golden", "small_grid.npz"))
    assert np.array_equal(got, golden[f"v_{family}_{int(clip)}"])


def _compare_runs(g, o, gf, of):
    gd, od = g.as_dict(), o.as_dict()
    assert gd == od, f"stats differ:\nGPU {gd}\nCPU {od}"
    assert np.array_equal(gf, of)


@pytest.mark.parametrize("raw", [False, True])
def test_small_grid_refinement(cpl, cplo, raw):
    cfg = abi.cfg_small_2param(raw=raw)
    g, gf = cpl.refine_run(cfg)
    o, of = cplo.refine_run(cfg)
    _compare_runs(g, o, gf, of)


def test_4param_refinement_reduced_grid(cpl, cplo):
    cfg = abi.cfg_4param(max_depth=6, grid=(16, 16, 16, 16))
    g, gf = cpl.refine_run(cfg)
    o, of = cplo.refine_run(cfg)
    _compare_runs(g, o, gf, of)
    assert g.total_evaluations > g.evaluations[0]  # refinement happened


def test_4param_full_grid_root_sample(cpl, cplo):
    """Full 64^4 grid on the GPU; the oracle re-runs a stride sample of the roots."""
    cfg = abi.cfg_4param(max_depth=6)
    g, gf = cpl.refine_run(cfg)
    stride = 4099
    o, of = cplo.refine_run(cfg, root_stride=stride)
    assert np.array_equal(gf[::stride], of)
    assert g.certified + g.unresolved == sum(
        g.evaluations[d] - (g.evaluations[d + 1] // 2 if d + 1 < g.rounds else 0) for d in range(g.rounds))


@pytest.mark.parametrize("family", range(6))
def test_reference_grid_odd_subdivisions(cpl, cplo, family):
    """Rows (la, r1) longer than a block, cell counts per row that 256 does not
    divide, and r3 counts below and above 256: every cell of the grid is
    evaluated once (same count and same first problematic la as the oracle)."""
    shapes = [(5, 7, 9, 300), (6, 7, 300, 3), (3, 5, 11, 37)] if abi.IS_3D[family] else [(7, 9, 1000, 1), (5, 3, 257, 1)]
    for subdiv in shapes:
        for outer in [(abi.LAMBDA_MIN, abi.LAMBDA_MAX), (0.55, 0.6)]:
            rc, first, ev, trap = cplo.first_problematic(family, outer, subdiv)
            if rc == 0:
                gfirst, gev, _ = cpl.first_problematic(family, outer, subdiv)
                assert (gfirst, gev) == (first, ev), (subdiv, outer)
            else:
                with pytest.raises(cpl.CplError) as ei:
                    cpl.first_problematic(family, outer, subdiv)
                assert ei.value.code == rc


# One launch per family over [lambda_min, lambda_max] on the reference's full
# grids (1e9 to 7e10 cells): first problematic la.lb and cells evaluated, as
# first measured with one flat cell index per thread (too large for the oracle;
# the truncated and odd-subdivision grids above are its parity checks).
FULL_GRID = {0: (0.5899609375, 1073741824), 1: (0.777060546875, 68719476736), 2: (0.5937890625, 1073741824),
             3: (0.6990625, 2147483648), 4: (0.7768212890625, 4294967296), 5: (0.8474023437500001, 2167963008)}


@pytest.mark.parametrize("family", range(6))
def test_reference_grid_full_first_launch(cpl, family):
    first, ev, _ = cpl.first_problematic(family, (abi.LAMBDA_MIN, abi.LAMBDA_MAX))
    assert (first, ev) == FULL_GRID[family]


@pytest.mark.parametrize("family", range(6))
def test_reference_grid_truncated(cpl, cplo, family):
    subdiv = (8, 8, 8, 4) if abi.IS_3D[family] else (16, 16, 16, 1)
    for outer in [(abi.LAMBDA_MIN, abi.LAMBDA_MAX), (0.5, 0.6), (0.7, 0.71)]:
        rc, first, ev, trap = cplo.first_problematic(family, outer, subdiv)
        if rc == 0:
            gfirst, gev, _ = cpl.first_problematic(family, outer, subdiv)
            assert (gfirst, gev) == (first, ev)
        else:
            with pytest.raises(cpl.CplError) as ei:
                cpl.first_problematic(family, outer, subdiv)
            assert ei.value.code == rc
