"""cpl_first_problematic_noalloc (fixed device scratch, no allocation) against
cpl_first_problematic_ex on the same launches: two calls in a row, the second
on a smaller grid, in every mode; sizes beyond the scratch are rejected."""
import pytest

from circlepacking_amd import abi

pytestmark = pytest.mark.gpu


def launch(cpl, noalloc, *args, **kw):
    try:
        first, info, text = cpl.first_problematic_ex(*args, noalloc=noalloc, text_cap=1 << 26, **kw)
        return 0, first, info, sorted(text.splitlines())
    except cpl.CplError as e:
        return e.code, e.first, e.info, sorted(e.text.splitlines())


CALLS = [  # (family, outer, subdiv): a larger launch, then a smaller one
    [(abi.START, (abi.LAMBDA_MIN, abi.LAMBDA_MAX), (32, 16, 16, 8)), (abi.START, (0.5, 0.6), (8, 8, 8, 4))],
    [(abi.END_2D, (abi.LAMBDA_MIN, abi.LAMBDA_MAX), (64, 32, 64, 1)), (abi.END_2D, (0.7, 0.71), (16, 8, 8, 1))],
    [(abi.MIDDLE, (0.55, 0.65), (16, 16, 16, 8)), (abi.END_3D, (0.5, 0.6), (6, 7, 9, 5))],
]


@pytest.mark.parametrize("calls", range(len(CALLS)))
@pytest.mark.parametrize("mode", ["early_exit", "exhaustive", "output"])
def test_noalloc_matches_the_allocating_entry_point(cpl, calls, mode):
    kw = {"exhaustive": mode == "exhaustive", "output_criticals": mode == "output"}
    for family, outer, subdiv in CALLS[calls]:
        print("case", family, outer, subdiv, mode, flush=True)
        a = launch(cpl, True, family, outer, subdiv, **kw)
        b = launch(cpl, False, family, outer, subdiv, **kw)
        assert a[0:2] == b[0:2]
        ia, ib = a[2], b[2]
        assert ia["trap_code"] == ib["trap_code"] and ia["fit_check_failed"] == ib["fit_check_failed"]
        if mode == "exhaustive":
            assert ia == ib
        if mode == "output":
            keys = ("cells_evaluated", "nofit_flags", "lines_found")
            assert [ia[k] for k in keys] == [ib[k] for k in keys]
            assert ib["lines_written"] == ib["lines_found"] == len(b[3])
            if ib["lines_found"] <= abi.NOALLOC_MAX_RECORDS:
                assert a[3] == b[3]
            else:  # the scratch keeps the first NOALLOC_MAX_RECORDS lines found
                assert ia["lines_written"] == len(a[3]) == abi.NOALLOC_MAX_RECORDS
                assert set(a[3]) <= set(b[3])


def test_noalloc_rejects_more_lambda_cells_than_its_scratch(cpl):
    with pytest.raises(cpl.CplError) as e:
        cpl.first_problematic_ex(abi.END_2D, (0.5, 0.6), (abi.NOALLOC_MAX_LA + 1, 2, 2, 1), noalloc=True)
    assert e.value.code == -2 and "exceed" in str(e.value)
    first, info, _ = cpl.first_problematic_ex(abi.END_2D, (0.5, 0.6), (abi.NOALLOC_MAX_LA, 2, 2, 1), noalloc=True)
    first2, _, _ = cpl.first_problematic_ex(abi.END_2D, (0.5, 0.6), (abi.NOALLOC_MAX_LA, 2, 2, 1))
    assert first == first2
