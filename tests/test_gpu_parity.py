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


def test_4param_full_grid_every_64th_root(cpl, cplo):
    """Full 64^4 grid, depth 6: the per-(root, family) outcome flags of every 64th
    root (262,144 roots, each refined to its leaves by the oracle) equal the GPU's."""
    cfg = abi.cfg_4param(max_depth=6)
    _, gf = cpl.refine_run(cfg)
    _, of = cplo.refine_run(cfg, root_stride=64)
    assert of.shape[0] == gf.shape[0] // 64
    assert np.array_equal(gf[::64], of)
    # the sample holds roots with criteria-decided leaves of both outcomes
    assert np.any(of & abi.RF_PROVEN) and np.any(of & abi.RF_CRITICAL)
