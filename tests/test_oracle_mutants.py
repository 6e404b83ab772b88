"""The oracle's pins have teeth: builds of the oracle with one planted mistake
each (a flipped weight, a wrong angle, a wrong constant, a wrong range bound,
a weakened case gate) fail the checks that pin the unmodified oracle.

* Every mistake in a criterion's formula is caught twice: by the exact-rational
  restatement written from the reference files (tests/exact_ref.py) and by the
  mpmath values of the area comparisons at point cells (tests/point_model.py).
* A mistake in the interval logic that leaves every point value unchanged (a
  gate's `definitely` weakened to `possibly`) is caught by the exact
  restatement only; a mistake in a range function by the exact range check.

The mistakes are planted in copies of oracle/cpl_oracle.hpp built into a
temporary directory; the oracle itself is not modified."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import criteria_checks as cc
from tests import exact_ref as X
from tests.cells import grid_cells, random_kernel_body_cells
from tests.test_oracle_criteria_exact import area_term_mismatches

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-frounding-math", "-shared"]

# name -> (family, kind, [(text in cpl_oracle.hpp, replacement)], reference line of the original)
MUTANTS = {
    "start_case1_weight": (0, "formula", [("double lb_disk_area = add_rd(lb_area1, mul_rd(0.5, lb_area2));",
                                           "double lb_disk_area = add_rd(lb_area1, lb_area2);")],
                           "start_configuration.cu:99"),
    "start_end_2d_reuse_angle": (1, "formula", [("IV reuse_angle = a12 + a2 - a1;", "IV reuse_angle = a12 + a1 - a2;")],
                                 "start_configuration.cu:175"),
    "start_end_3d_a13_pair": (2, "formula", [("IV a13 = collision_angle_outer_outer(r1, r2);",
                                              "IV a13 = collision_angle_outer_outer(r1, r3);")],
                              "start_configuration.cu:232"),
    "middle_case_b_angle": (3, "formula", [("double ub_area_used = mul_ru(ub_ratio, mul_ru(0.5, c13.ub));",
                                            "double ub_area_used = mul_ru(ub_ratio, mul_ru(0.5, c12.ub));")],
                            "middle_configuration.cu:110,144"),
    "end_2d_inner_rho": (4, "formula", [("IV rho = (1.0 - 2.0 * r2).square() - la.square();",
                                         "IV rho = (1.0 - r2).square() - la.square();")],
                         "end_configuration.cu:78"),
    "critical_ratio": (5, "formula", [("static const double critical_ratio = 0.5642;",
                                       "static const double critical_ratio = 0.5652;")], "values.hpp:28"),
    "end_3d_outer_gate": (5, "logic", [("IV a13 = collision_angle_outer_outer(r1, r3);\n"
                                        "  if (definitely(a12 + a23 >= a13)) return true;",
                                        "IV a13 = collision_angle_outer_outer(r1, r3);\n"
                                        "  if (possibly(a12 + a23 >= a13)) return true;")],
                          "end_configuration.cu:188"),
    "end_r2_range": (4, "range", [("double l = mul_rd(0.5, add_rd(1.0, -add_ru(la.ub, mul_ru(2.0, r1.ub))));",
                                   "double l = mul_rd(0.5, add_rd(1.0, -add_ru(la.ub, mul_ru(2.0, r1.lb))));")],
                     "end_configuration.cu:66"),
    "r1_range_lane_width": (0, "range", [("  double lb_by_lane_width = (0.25 * (1.0 - la)).lb;\n  return IV(lb_by_lane_width,",
                                          "  double lb_by_lane_width = (0.2 * (1.0 - la)).lb;\n  return IV(lb_by_lane_width,")],
                            "start_configuration.cu:64"),
}

_P_D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    """Builds every mutant in parallel; name -> loaded library."""
    base = tmp_path_factory.mktemp("mutants")
    src = open(os.path.join(ORACLE, "cpl_oracle.hpp")).read()
    procs = {}
    for name, (_, _, edits, _) in MUTANTS.items():
        d = base / name / "oracle"
        d.mkdir(parents=True)
        (base / name / "include").mkdir()
        shutil.copy(os.path.join(ROOT, "include", "cpl.h"), base / name / "include" / "cpl.h")
        shutil.copy(os.path.join(ORACLE, "cpl_oracle_capi.cpp"), d / "cpl_oracle_capi.cpp")
        text = src
        for old, new in edits:
            assert text.count(old) == 1, (name, old)  # the planted mistake must still apply
            text = text.replace(old, new)
        (d / "cpl_oracle.hpp").write_text(text)
        so = d / "liboracle_mut.so"
        procs[name] = (so, subprocess.Popen(["g++"] + CXXFLAGS + ["-o", str(so), str(d / "cpl_oracle_capi.cpp"),
                                                                 "-lpthread"], stderr=subprocess.PIPE))
    libs = {}
    for name, (so, p) in procs.items():
        _, err = p.communicate()
        assert p.returncode == 0, err.decode()[-2000:]
        L = ctypes.CDLL(str(so))
        L.cplo_eval_kernel_body.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint8)]
        L.cplo_area_terms.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, _P_D, ctypes.POINTER(ctypes.c_int)]
        L.cplo_ranges.argtypes = [ctypes.c_int, _P_D, _P_D, ctypes.c_size_t]
        L.cplo_prim.argtypes = [ctypes.c_int, _P_D, _P_D, _P_D, ctypes.c_size_t]
        libs[name] = L
    return libs


@pytest.fixture(scope="module")
def ref(cplo):
    return X.Reference(lambda x: float(cplo.prim("asin_rn", np.array([x]))[0]))


def _evaluator(L):
    def ev(family, cells):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        out = np.empty(cells.shape[0], dtype=np.uint8)
        assert L.cplo_eval_kernel_body(family, cells.ctypes.data_as(_P_D), cells.shape[0],
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))) == 0
        return out
    return ev


def exact_mismatches(evaluator, area_terms, ref, cplo, family):
    """Cells where the evaluator's verdict byte or area comparisons differ from
    the exact restatement's: a stratified sample of random boxes and
    reference-grid cells."""
    pool = random_kernel_body_cells(900 + family, 8000, family)
    pb = cplo.eval_kernel_body(family, pool)
    pick = []
    rng = np.random.default_rng(family)
    for v in np.unique(pb):
        idx = np.nonzero(pb == v)[0]
        pick.extend(rng.choice(idx, min(40, idx.size), replace=False).tolist())
    cells = np.concatenate([pool[sorted(pick)], grid_cells(ref, family, 200, 5, (16, 16, 16, 8)),
                            grid_cells(ref, family, 200, 6, (64, 32, 32, 8))])
    got = evaluator(family, cells)
    n = 0
    for c, b in zip(cells, got):
        want = ref.kernel_body(family, X.IV(c[0], c[1]), X.IV(c[2], c[3]), X.IV(c[4], c[5]), X.IV(c[6], c[7]))
        n += int(b) != want
    return n + len(area_term_mismatches(area_terms, ref, family, cells))


def range_mismatches(L, ref, family):
    rng = np.random.default_rng(7)
    n = 300
    la0 = rng.uniform(0.5, 0.98, n)
    lane = 1.0 - la0
    r1 = lane * rng.uniform(0.2, 0.5, n)
    r2 = r1 * rng.uniform(0.5, 1.0, n)
    recs = np.ascontiguousarray(np.stack([la0, la0 + 1e-3, r1, r1 + lane * 1e-2, r2, r2 + lane * 1e-3], 1))
    out = np.empty_like(recs)
    assert L.cplo_ranges(family, recs.ctypes.data_as(_P_D), out.ctypes.data_as(_P_D), n) == 0
    bad = 0
    for rec, row in zip(recs, out):
        la, ra, rb = X.IV(rec[0], rec[1]), X.IV(rec[2], rec[3]), X.IV(rec[4], rec[5])
        a, b = ref.compute_r1_range(la), ref.compute_r2_range(family, la, ra)
        c = ref.compute_r3_range(family, la, ra, rb) if family in X.THREE_D else X.IV(0.0, 0.0)
        bad += not np.array_equal(np.array([a.lb, a.ub, b.lb, b.ub, c.lb, c.ub]), row)
    return bad


def _area_terms(L):
    def area_terms(f, cells):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        n = cells.shape[0]
        out = np.zeros((n, 16))
        cnt = np.zeros(n, dtype=np.int32)
        assert L.cplo_area_terms(f, cells.ctypes.data_as(_P_D), n, out.ctypes.data_as(_P_D),
                                 cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
        return out.reshape(n, 8, 2), cnt
    return area_terms


def mpmath_mismatches(L, family, lbrsp):
    points = np.concatenate([cc.domain_points(family, 300, 11, lbrsp), cc.wide_points(family, 300, 12)])
    bad, compared = cc.term_mismatches(_area_terms(L), family, points)
    assert compared > 100
    return len(bad)


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistake_is_caught(mutant_libs, ref, cplo, name):
    family, kind, _, where = MUTANTS[name]
    L = mutant_libs[name]
    lbrsp = lambda la: float(cplo.prim("lbrsp", np.array([la]))[0])
    if kind == "range":
        assert range_mismatches(L, ref, family) > 0, f"{name} ({where}) not caught by the exact range check"
        return
    assert exact_mismatches(_evaluator(L), _area_terms(L), ref, cplo, family) > 0, \
        f"{name} ({where}) not caught by the exact check"
    n_mp = mpmath_mismatches(L, family, lbrsp)
    if kind == "formula":
        assert n_mp > 0, f"{name} ({where}) not caught by the mpmath area terms"
    else:
        assert n_mp == 0  # interval logic only: the point values are unchanged


def test_unmodified_oracle_passes_the_same_checks(ref, cplo):
    lbrsp = lambda la: float(cplo.prim("lbrsp", np.array([la]))[0])
    for family in range(6):
        assert exact_mismatches(cplo.eval_kernel_body, cplo.area_terms, ref, cplo, family) == 0
        assert range_mismatches(cplo.lib(), ref, family) == 0
        assert mpmath_mismatches(cplo.lib(), family, lbrsp) == 0
