"""Without a GPU the library loads and every entry point fails cleanly with
CPL_ERR_NO_DEVICE (no crash, a message in cpl_last_error)."""
import ctypes

import pytest


def test_calls_fail_cleanly_without_a_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import circlepacking_amd as cpl
    from circlepacking_amd import abi

    lib = cpl.lib()
    assert lib.cpl_init() == -6
    assert b"device" in lib.cpl_last_error().lower()
    st = abi.RefineStats()
    cfg = abi.cfg_4param(grid=(4, 4, 4, 4), max_depth=1)
    assert lib.cpl_refine_run(ctypes.byref(cfg), ctypes.byref(st), None) == -6
    assert lib.cpl_shutdown() == 0
