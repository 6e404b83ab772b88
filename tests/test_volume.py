"""Leaf volumes of a refinement run (the measure the learned split policy is
compared by): every root cell is partitioned into its leaves, so certified and
unresolved volume add up to the root box for each family, and one more level
of refinement never increases the unresolved volume."""
import pytest

from circlepacking_amd import abi


@pytest.mark.parametrize("cfg", [abi.cfg_4param(grid=(8, 8, 8, 8), max_depth=4),
                                 abi.cfg_small_2param(raw=True, n=48, max_depth=3),
                                 abi.cfg_small_2param(raw=False, n=48, max_depth=3)])
def test_leaf_volumes_partition_the_box(cplo, cfg):
    st, _ = cplo.refine_run(cfg)
    cert, unres = st.certified_volume(cfg), st.unresolved_volume(cfg)
    for f in cert:
        assert cert[f] + unres[f] == pytest.approx(1.0, abs=1e-12), (f, cert[f], unres[f])
        assert 0.0 <= unres[f] <= 1.0


def test_unresolved_volume_never_grows_with_depth(cplo):
    prev = None
    for depth in range(0, 5):
        cfg = abi.cfg_4param(grid=(6, 6, 6, 6), max_depth=depth)
        st, _ = cplo.refine_run(cfg)
        vol = st.unresolved_volume(cfg)
        if prev is not None:
            for f in vol:
                assert vol[f] <= prev[f] + 1e-15, (depth, f, vol[f], prev[f])
        prev = vol
