"""The C-ABI library loads and exports every function include/cpl.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cpl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(cpl_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("cpl_first_problematic", "cpl_eval_cells", "cpl_refine_run", "cpl_last_error", "cpl_init"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import circlepacking_amd

    lib = ctypes.CDLL(circlepacking_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    from circlepacking_amd import abi

    assert ctypes.sizeof(abi.Iv) == 16
    assert ctypes.sizeof(abi.RefineCfg) == 4 * 4 + 8 * 8 + 4 * 5 + 4 * 7
    # stats: 2 ints, 16 evaluations, 16x8 verdicts, 2x16 digests, 16x6x8 family verdicts,
    # 16x6 launch seconds, 3 counters, 2 doubles, peak
    expect = 8 + 8 * (16 + 128 + 32 + 16 * 48) + 8 * 96 + 8 * 3 + 8 * 2 + 8
    assert ctypes.sizeof(abi.RefineStats) == expect
