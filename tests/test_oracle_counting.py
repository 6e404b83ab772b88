"""The oracle's flop-counting build (the algorithmic flops behind the roofline
figure of bench.py): counting changes no result, and the per-evaluation counts
have the structure the criteria imply."""
import numpy as np

from circlepacking_amd import abi
from oracle import cplo


def _key(st):
    return ([[st.family_verdicts[d][f][v] for v in range(8)] for d in range(st.rounds) for f in range(6)],
            list(st.digest[:st.rounds]), list(st.verdict_digest[:st.rounds]), st.certified, st.unresolved)


def test_counting_changes_no_result_and_counts_follow_the_criteria():
    cfg = abi.cfg_4param(grid=(16, 16, 16, 16), max_depth=5)
    st0, _ = cplo.refine_run(cfg, want_flags=False, nthreads=4)
    st1, _, flops = cplo.refine_run(cfg, want_flags=False, nthreads=4, counting=True)
    assert _key(st0) == _key(st1)
    counts = np.zeros((6, 8))
    for d in range(st1.rounds):
        counts += np.array([[st1.family_verdicts[d][f][v] for v in range(8)] for f in range(6)], dtype=np.float64)
    per_eval = np.where(counts > 0, flops / np.maximum(counts, 1), 0.0)
    for f in cfg.families():
        # out of domain: the admissible-range arithmetic only (two interval
        # subtractions and two scalings settle the first range test)
        if counts[f][abi.V_OOD]:
            assert 8 <= per_eval[f][abi.V_OOD] < 40, (f, per_eval[f][abi.V_OOD])
        # a verdict from the criteria evaluates at least three collision angles
        for v in (abi.V_PROVEN, abi.V_CRITICAL):
            if counts[f][v]:
                assert 300 < per_eval[f][v] < 4000, (f, v, per_eval[f][v])
    # middle and end 3d run a second criterion when the first holds: a proven
    # cell costs more than a cell that fails the first criterion
    for f in (abi.MIDDLE, abi.END_3D):
        if counts[f][abi.V_PROVEN] and counts[f][abi.V_CRITICAL]:
            assert per_eval[f][abi.V_PROVEN] > per_eval[f][abi.V_CRITICAL]
