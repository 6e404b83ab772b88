"""Mode R on the GPU: one launch of a reference kernel over a lambda interval on
the reference grid (cpl_first_problematic_ex), against the oracle's
reference-grid search: first problematic la, failure rule and trap code, cells
evaluated (exhaustive launches), and the printed lines of output mode."""
import pytest

from circlepacking_amd import abi

pytestmark = pytest.mark.gpu

OUTERS = [(abi.LAMBDA_MIN, abi.LAMBDA_MAX), (0.5, 0.6), (0.55, 0.6), (0.7, 0.71)]


def gpu(cpl, family, outer, subdiv, output=False, exhaustive=False, text_cap=1 << 22):
    """-> (rc, first, info, text) of one launch."""
    try:
        first, info, text = cpl.first_problematic_ex(family, outer, subdiv, output, exhaustive, text_cap)
        return 0, first, info, text
    except cpl.CplError as e:
        return e.code, e.first, e.info, e.text


def shapes(family):
    """Truncated grids, rows longer than a block, cell counts per row that 64
    does not divide, r3 counts below and above 64."""
    if abi.IS_3D[family]:
        return [(8, 8, 8, 4), (5, 7, 9, 300), (6, 7, 300, 3), (3, 5, 11, 37)]
    return [(16, 16, 16, 1), (7, 9, 1000, 1), (5, 3, 257, 1)]


def check_against_oracle(cpl, cplo, family, outer, subdiv):
    o = cplo.mode_r(family, outer, subdiv)
    for exhaustive in (True, False):
        rc, first, info, _ = gpu(cpl, family, outer, subdiv, exhaustive=exhaustive)
        case = (family, outer, subdiv, exhaustive)
        print("case", case, flush=True)
        assert (rc, first) == (o["rc"], o["first"]), (case, rc, first, o)
        assert info["trap_code"] == o["trap_code"], case
        if exhaustive:
            assert info["cells_evaluated"] == o["evaluated"] and info["rows_skipped"] == 0, (case, info, o)
            assert info["nofit_flags"] == o["nofit_flags"], case
        else:
            assert info["cells_evaluated"] <= o["evaluated"], case
    return o


@pytest.mark.parametrize("family", range(6))
def test_truncated_and_odd_grids(cpl, cplo, family):
    skipped_somewhere = False
    for subdiv in shapes(family):
        for outer in OUTERS:
            o = check_against_oracle(cpl, cplo, family, outer, subdiv)
            if o["rc"] == 0 and o["first"] < outer[1]:
                _, _, info, _ = gpu(cpl, family, outer, subdiv)
                skipped_somewhere |= info["rows_skipped"] > 0 or info["cells_evaluated"] < o["evaluated"]
    assert skipped_somewhere


# One launch per family over [lambda_min, lambda_max] on the reference's full
# grids (1e9 to 7e10 cells, beyond the oracle): first problematic la.lb and
# cells evaluated by an exhaustive launch, as first measured by the kernel that
# evaluated every cell (the truncated and odd grids above are the oracle checks).
FULL_GRID = {0: (0.5899609375, 1073741824), 1: (0.777060546875, 68719476736), 2: (0.5937890625, 1073741824),
             3: (0.6990625, 2147483648), 4: (0.7768212890625, 4294967296), 5: (0.8474023437500001, 2167963008)}


@pytest.mark.parametrize("family", range(6))
def test_full_grid_first_launch(cpl, family):
    outer = (abi.LAMBDA_MIN, abi.LAMBDA_MAX)
    rc, first, info, _ = gpu(cpl, family, outer, None, exhaustive=True)
    assert rc == 0 and (first, info["cells_evaluated"]) == FULL_GRID[family]
    rc, first2, info2, _ = gpu(cpl, family, outer, None)
    assert rc == 0 and first2 == first
    # the early exit skips the rows at or above the minimum
    assert info2["rows_skipped"] > 0 and info2["cells_evaluated"] < info["cells_evaluated"]


HEADS = {
    (0, 1): "Critical interval: Start configuration, case 1: ",
    (0, 2): "Critical interval: Start configuration, case 2: ",
    (0, 3): "Start configuration: Circle possibly does not fit - ",
    (1, 1): "Critical interval: Start & end configuration, 2 disks: ",
    (2, 1): "Start & end configuration: Critical interval - ",
    (3, 1): "Critical interval: Middle configuration, case 1.1: ",
    (3, 2): "Critical interval: Middle configuration, case 2.1: ",
    (4, 1): "End configuration (2 disks): Critical interval with r1 on the inside - ",
    (4, 2): "End configuration (2 disks): Critical interval with r1 on the outside - ",
    (5, 1): "End configuration (3 disks): Critical interval with r1 on the inside - ",
    (5, 2): "End configuration (3 disks): Critical interval with r1 on the outside - ",
}


def reference_line(family, rec):
    """The printf line of the reference for an oracle record (start_configuration.cu:132,147,158,207,289;
    middle_configuration.cu:182,193; end_configuration.cu:124,135,245,256)."""
    k = int(rec[0])
    g = ["%.19g" % x for x in rec[1:]]
    body = f"la = [{g[0]},{g[1]}], r1 = [{g[2]},{g[3]}], r2 = [{g[4]},{g[5]}]"
    if abi.IS_3D[family]:
        body += f", r3 = [{g[6]},{g[7]}]"
    return HEADS[(family, k)] + body


@pytest.mark.parametrize("family", range(6))
def test_output_mode_lines(cpl, cplo, family):
    lines = 0
    small = [(8, 8, 8, 4), (3, 5, 11, 37)] if abi.IS_3D[family] else [(16, 16, 16, 1), (5, 3, 257, 1)]
    for subdiv in small:
        for outer in OUTERS:
            o = cplo.mode_r(family, outer, subdiv, output_criticals=True, rec_cap=1 << 20)
            rc, first, info, text = gpu(cpl, family, outer, subdiv, output=True)
            case = (family, outer, subdiv)
            print("case", case, flush=True)
            assert (rc, first, info["trap_code"]) == (o["rc"], o["first"], o["trap_code"]), case
            assert info["cells_evaluated"] == o["evaluated"] and info["nofit_flags"] == o["nofit_flags"], case
            want = sorted(reference_line(family, r) for r in o["records"])
            got = sorted(text.splitlines())
            assert info["lines_found"] == o["n_records"] == len(want) == info["lines_written"], case
            assert got == want, case
            assert info["text_bytes_needed"] == sum(len(l) + 1 for l in want) + 1
            lines += len(want)
    assert lines > 0


def test_output_mode_reports_truncation(cpl, cplo):
    family, outer, subdiv = abi.END_2D, (abi.LAMBDA_MIN, abi.LAMBDA_MAX), (16, 16, 16, 1)
    o = cplo.mode_r(family, outer, subdiv, output_criticals=True)
    assert o["n_records"] > 3
    rc, _, info, text = gpu(cpl, family, outer, subdiv, output=True, text_cap=300)
    assert info["lines_found"] == o["n_records"] and info["lines_written"] == len(text.splitlines()) <= 1
    assert info["text_bytes_needed"] > 300
