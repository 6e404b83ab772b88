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


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors against the C compiler's layout of include/cpl.h."""
    import subprocess

    from circlepacking_amd import abi

    fields = {
        "cpl_refine_cfg": (abi.RefineCfg, ["grid", "box_lb", "box_ub", "family_mask", "clip", "max_depth",
                                           "split_mode", "split_axes", "options", "block", "cells_per_thread",
                                           "prefilter_depths", "reserved"]),
        "cpl_refine_stats": (abi.RefineStats, [f[0] for f in abi.RefineStats._fields_]),
        "cpl_tune_point": (abi.TunePoint, ["block", "cells_per_thread", "seconds"]),
        "cpl_iv": (abi.Iv, ["lb", "ub"]),
        "cpl_mode_r_info": (abi.ModeRInfo, [f[0] for f in abi.ModeRInfo._fields_]),
    }
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "cpl.h"', "int main(void) {"]
    for ctype, (_, names) in fields.items():
        src.append(f'  printf("{ctype} %zu\\n", sizeof({ctype}));')
        for n in names:
            src.append(f'  printf("{ctype}.{n} %zu\\n", offsetof({ctype}, {n}));')
    src.append("  return 0;\n}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.splitlines())
    for ctype, (py, names) in fields.items():
        assert int(got[ctype]) == ctypes.sizeof(py), ctype
        for n in names:
            assert int(got[f"{ctype}.{n}"]) == getattr(py, n).offset, (ctype, n)
