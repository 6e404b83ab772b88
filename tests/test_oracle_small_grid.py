"""The oracle reproduces the committed golden verdicts of the 2-parameter grid,
and its synthetic refinement obeys the structural invariants of bisection."""
import json
import os

import numpy as np
import pytest

from circlepacking_amd import abi
from tests.cells import c1_grid_soa

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_grid.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("family", [abi.START_END_2D, abi.END_2D])
@pytest.mark.parametrize("clip", [True, False])
def test_cell_verdicts_match_golden(cplo, golden, family, clip):
    got = cplo.eval_cells(family, c1_grid_soa(256), clip)
    assert np.array_equal(got, golden[f"v_{family}_{int(clip)}"])


@pytest.mark.parametrize("raw", [False, True])
def test_refinement_matches_golden(cplo, golden, raw):
    st, flags = cplo.refine_run(abi.cfg_small_2param(raw=raw))
    runs = json.loads(bytes(golden["runs_json"]).decode())
    assert st.as_dict() == runs[f"raw{int(raw)}"]
    assert np.array_equal(flags, golden[f"flags_raw{int(raw)}"])


def test_raw_grid_exercises_every_verdict(golden):
    kinds = set()
    for k in golden.files:
        if k.startswith("v_"):
            kinds |= set((golden[k] & 7).tolist())
    assert {abi.V_PROVEN, abi.V_NOFIT, abi.V_EMPTY, abi.V_OOD, abi.V_TRAP} <= kinds


def test_bisection_invariants(cplo):
    """Each critical, non-final item yields exactly two children (mode 0) or four
    (2x2 mode), so evaluations[d+1] = k * critical[d]."""
    cfg = abi.cfg_4param(max_depth=5, grid=(8, 8, 8, 8))
    st, flags = cplo.refine_run(cfg)
    for d in range(st.rounds - 1):
        assert st.evaluations[d + 1] == 2 * st.verdicts[d][abi.V_CRITICAL]
    cfg2 = abi.cfg_small_2param(raw=True, n=64)
    st2, _ = cplo.refine_run(cfg2)
    assert st2.verdicts[0][abi.V_CRITICAL] > 0 and st2.rounds == 3
    for d in range(st2.rounds - 1):
        assert st2.evaluations[d + 1] == 4 * st2.verdicts[d][abi.V_CRITICAL]
    # leaves = certified + unresolved
    leaves = sum(st.evaluations[d] - (st.evaluations[d + 1] // 2 if d + 1 < st.rounds else 0) for d in range(st.rounds))
    assert leaves == st.certified + st.unresolved
    # root flags: every root has at least one leaf outcome per family
    assert np.all(flags != 0)


def test_root_cell_r3_position_equals_nested_get_subinterval(cplo):
    """For a root cell, the r3 position interval [k/n, (k+1)/n] selects exactly
    get_subinterval(range, k, n) (subintervals.hpp:31-42): checked through the
    reference kernel body on explicit cells vs the synthetic evaluation."""
    from tests.cells import c1_grid_soa as _  # noqa: F401
    rng = np.random.default_rng(9)
    n = 16
    cells, soa = [], []
    for _ in range(200):
        la = rng.uniform(0.5, 0.8)
        lane = 1 - la
        r1 = lane * rng.uniform(0.28, 0.3)
        r2 = r1 * rng.uniform(0.95, 1.0)
        k = int(rng.integers(0, n))
        lo = 0.25 * lane  # quarter-lane bound of the middle family's r3 range (point la)
        import math
        hi = r2
        diff = (hi - lo) / n
        from fractions import Fraction
        l = float(Fraction(diff) * k + Fraction(lo))
        u = hi if k == n - 1 else float(Fraction(diff) * (k + 1) + Fraction(lo))
        cells.append([la, la, r1, r1, r2, r2, l, u])
        soa.append([la, la, r1, r1, r2, r2, k / n, (k + 1) / n])
    cells = np.array(cells)
    soa = np.array(soa).T.copy()
    want = cplo.eval_kernel_body(abi.MIDDLE, cells)
    got = cplo.eval_cells(abi.MIDDLE, soa, clip=False)
    # (0.25*(1-la)).lb rounds 1-la down; for these la the product is exact, so the ranges agree
    assert np.array_equal(got, want)
