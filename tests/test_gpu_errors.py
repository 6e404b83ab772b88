"""Error paths of the C ABI on the GPU: invalid configurations and arguments
return CPL_ERR_ARG, an outgrown worklist returns CPL_ERR_CAPACITY in every
execution form, and the library keeps working afterwards."""
import numpy as np
import pytest

from circlepacking_amd import abi

pytestmark = pytest.mark.gpu


def _code(cpl, fn):
    with pytest.raises(cpl.CplError) as ei:
        fn()
    assert cpl.lib().cpl_last_error()  # a message is always set
    return ei.value.code


def test_invalid_configurations(cpl):
    base = abi.cfg_4param(grid=(8, 8, 8, 8), max_depth=2)
    bad = []
    c = abi.with_options(base, 0); c.grid[1] = 0; bad.append(c)
    c = abi.with_options(base, 0); c.split_axes = 0; bad.append(c)
    c = abi.with_options(base, 0); c.split_axes = 0x10; bad.append(c)
    c = abi.with_options(base, 0); c.max_depth = abi.MAX_ROUNDS; bad.append(c)
    c = abi.with_options(base, 0); c.family_mask = 0; bad.append(c)
    bad.append(abi.with_options(base, abi.OPT_GRAPH, block=100))
    bad.append(abi.with_options(base, abi.OPT_GRAPH, block=2048))
    bad.append(abi.with_options(base, abi.OPT_HIER_ROOTS))  # needs a batched form
    bad.append(abi.with_options(base, 1 << 20))              # unknown option bit
    for cfg in bad:
        assert _code(cpl, lambda: cpl.refine_run(cfg)) == -2
    # still usable
    s, f = cpl.refine_run(base)
    assert s.total_evaluations > 0


def test_invalid_arguments(cpl):
    lib = cpl.lib()
    assert lib.cpl_debug_prim(99, None, None, None, 0) == -2
    iv = np.zeros((4, 2))
    assert _code(cpl, lambda: cpl.interval_op("add", iv)) == -2  # binary op without b
    assert lib.cpl_interval_op(16, None, None, None, None, 0) == -2
    assert _code(cpl, lambda: cpl.first_problematic(7, (0.5, 0.6))) == -2
    assert _code(cpl, lambda: cpl.eval_kernel_body(6, np.zeros((1, 8)))) == -2


@pytest.mark.parametrize("opts", [0, abi.OPT_GRAPH | abi.OPT_HIER_ROOTS, abi.OPT_PERSISTENT | abi.OPT_HIER_ROOTS])
def test_capacity_exceeded_is_reported(cpl, monkeypatch, opts):
    base = abi.with_options(abi.cfg_4param(grid=(16, 16, 16, 16), max_depth=4), opts)
    ref = cpl.refine_run(abi.with_options(base, 0))
    # the persistent form keeps children in per-wave stacks, so only its root
    # list is bounded by the worklist capacity
    monkeypatch.setenv("CPL_WORKLIST_CAP", "256" if opts & abi.OPT_PERSISTENT else "4096")
    assert _code(cpl, lambda: cpl.refine_run(base)) == -5
    monkeypatch.delenv("CPL_WORKLIST_CAP")
    s, f = cpl.refine_run(base)  # recovers with the default capacity
    assert s.as_dict() == ref[0].as_dict() and np.array_equal(f, ref[1])
