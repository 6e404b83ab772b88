"""The oracle's second-check-first order (its stand-in for a proposal that
always votes for the second check of the middle, end 2-disk and end 3-disk
families): verdict kinds, counts, item digests and root flags equal those of
the reference order; only the failing-criterion details change, and only in
refinement rounds (the root round runs without a proposal)."""
import numpy as np

from circlepacking_amd import abi
from oracle import cplo

KIND_FIELDS = ("rounds", "evaluations", "verdicts", "digest", "family_verdicts", "total_evaluations", "certified",
               "unresolved")


def test_second_check_first_keeps_kinds_and_changes_details():
    base = abi.cfg_4param(grid=(12, 12, 12, 12), families=(abi.MIDDLE, abi.END_3D), max_depth=4)
    a, fa = cplo.refine_run(base)
    b, fb = cplo.refine_run(abi.with_options(base, abi.OPT_LEARNED_ORDER))
    da, db = a.as_dict(), b.as_dict()
    for k in KIND_FIELDS:
        assert da[k] == db[k], k
    assert np.array_equal(fa, fb)
    assert da["verdict_digest"][0] == db["verdict_digest"][0]
    assert da["verdict_digest"][1:] != db["verdict_digest"][1:]


def test_order_has_no_effect_on_single_check_families():
    base = abi.cfg_4param(grid=(12, 12, 12, 12), families=(abi.START, abi.START_END_3D), max_depth=3)
    a, fa = cplo.refine_run(base)
    b, fb = cplo.refine_run(abi.with_options(base, abi.OPT_LEARNED_ORDER))
    assert a.as_dict() == b.as_dict()
    assert np.array_equal(fa, fb)
