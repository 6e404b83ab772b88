"""Writes tests/golden/small_grid.npz: per-cell verdict bytes of the
2-parameter grid (lambda = 0.5, r1, r2 in [0.05, 0.5], 256 x 256 cells) for the
start&end 2-disk and end 2-disk families, clipped to the admissible ranges and
raw, computed by the CPU oracle; plus the oracle's refinement statistics for
the same grid (2x2 splits to depth 2).  Run from the repository root:
    python tests/golden/gen_small_grid.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from circlepacking_amd import abi  # noqa: E402
from oracle import cplo  # noqa: E402
from tests.cells import c1_grid_soa  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_grid.npz")


def main():
    soa = c1_grid_soa(256)
    out = {}
    for f in (abi.START_END_2D, abi.END_2D):
        for clip in (True, False):
            out[f"v_{f}_{int(clip)}"] = cplo.eval_cells(f, soa, clip)
    runs = {}
    for raw in (False, True):
        st, flags = cplo.refine_run(abi.cfg_small_2param(raw=raw))
        runs[f"raw{int(raw)}"] = st.as_dict()
        out[f"flags_raw{int(raw)}"] = flags
    out["runs_json"] = np.frombuffer(json.dumps(runs).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
