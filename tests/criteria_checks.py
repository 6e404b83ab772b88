"""Checks of an evaluator of the kernel body against the real-valued criteria
of tests/point_model.py — TEST INFRASTRUCTURE.

An evaluator is a function (family, cells (n, 8): la, r1_, r2_, r3 as lb/ub
pairs) -> verdict bytes; the oracle's eval_kernel_body is one, and so is a
deliberately broken build of it (the mutation tests).  Two checks:

* soundness: every PROVEN cell satisfies the real criterion at random interior
  points (r1 >= r2 >= r3, the order the tightening of the kernel bodies
  assumes);
* agreement on point cells: a degenerate cell [p, p] is PROVEN exactly when the
  real criterion holds at p, except within a rounding-sized distance of a
  decision boundary (the interval evaluation at a point is exact up to a few
  ulps per operation);
* area terms on point cells: the two sides of every area comparison the
  evaluator makes (lower bound of the disk area, upper bound of c x used area;
  the oracle's area probe) enclose the real values at p and lie within a
  relative REL of them."""
import numpy as np

from circlepacking_amd import abi
from tests import point_model as pm

TOL = 1e-9  # excluded band around a decision boundary (relative to the lane area / in radians)
REL = 1e-8  # rounding of the interval evaluation at a point (cancellation near lambda = 0.99), relative


def lane_bounds(family, la, r1=None, r2=None):
    """Lower bounds of r2 (given r1) and r3 (given r2) of the family's
    admissible ranges (start_configuration.cu:68-87, middle_configuration.cu:
    76-84, end_configuration.cu:65-69, 148-152), in floating point (sampling only)."""
    q = 0.25 * (1 - la)
    if r2 is None:
        if family == abi.START_END_2D:
            return 0.5 * (1 - la - 2 * r1)
        if family in (abi.END_2D, abi.END_3D):
            return max(0.0, 0.5 * (1 - la - 2 * r1))
        return q
    if family == abi.START_END_3D:
        return 0.5 * (1 - la - 2 * r2)
    if family == abi.END_3D:
        return max(0.0, 0.5 * (1 - la - 2 * r2))
    return q


def domain_points(family, n, seed, lbrsp):
    """n points (la, r1, r2, r3) of the family's admissible domain, r1 >= r2 >= r3.
    lbrsp: the single-placement bound as a function of la (upper end of r1)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        la = rng.uniform(0.5, 0.99)
        lo1, hi1 = 0.25 * (1 - la), min(0.5 * (1 - la), lbrsp(la))
        if hi1 <= lo1:
            continue
        r1 = rng.uniform(lo1, hi1)
        lo2 = lane_bounds(family, la, r1)
        if lo2 >= r1:
            continue
        r2 = rng.uniform(lo2, r1)
        r3 = 0.0
        if abi.IS_3D[family]:
            lo3 = lane_bounds(family, la, r1, r2)
            if lo3 >= r2:
                continue
            r3 = rng.uniform(lo3, r2)
        out.append((la, r1, r2, r3))
    return np.array(out)


def point_cells(points):
    return np.repeat(points, 2, axis=1)


def model(family, p, physical=True):
    return pm.evaluate(family, p[0], p[1], p[2], p[3] if abi.IS_3D[family] else None, physical)


def point_disagreements(evaluator, family, points):
    """-> (disagreeing points, counts of PROVEN / CRITICAL among the compared ones)."""
    vb = evaluator(family, point_cells(points))
    bad, proven, critical = [], 0, 0
    for p, b in zip(points, vb):
        v = b & 7
        if v not in (abi.V_PROVEN, abi.V_CRITICAL):
            continue
        holds, margin, _ = model(family, p)
        if margin < TOL:
            continue
        if holds != (v == abi.V_PROVEN):
            bad.append((p.tolist(), int(b), holds, float(margin)))
        proven += v == abi.V_PROVEN
        critical += v == abi.V_CRITICAL
    return bad, proven, critical


def in_domain(family, p, lbrsp):
    """The point lies in the family's admissible domain (floating point)."""
    la, r1, r2, r3 = p
    if not (0.5 <= la <= 0.99 and 0.25 * (1 - la) <= r1 <= min(0.5 * (1 - la), lbrsp(la))):
        return False
    if not (lane_bounds(family, la, r1) <= r2 <= r1):
        return False
    return not abi.IS_3D[family] or lane_bounds(family, la, r1, r2) <= r3 <= r2


def interior_points(cell, family, k, rng):
    """Up to k points of the tightened cell with r1 >= r2 >= r3 (rejection)."""
    la, r1, r2, r3 = cell[0:2], cell[2:4].copy(), cell[4:6].copy(), cell[6:8]
    if abi.IS_3D[family]:
        r2[0] = max(r2[0], r3[0])
    r1[0] = max(r1[0], r2[0])
    pts = []
    for _ in range(50 * k):
        if len(pts) == k:
            break
        p3 = rng.uniform(r3[0], r3[1]) if abi.IS_3D[family] else 0.0
        p2 = rng.uniform(r2[0], r2[1])
        p1 = rng.uniform(r1[0], r1[1])
        if p1 >= p2 and (not abi.IS_3D[family] or p2 >= p3):
            pts.append((rng.uniform(la[0], la[1]), p1, p2, p3))
    return pts


def soundness_violations(evaluator, family, cells, seed, lbrsp, k=4):
    """-> (points of PROVEN cells in the admissible domain where the real
    criterion fails, points checked).  Outside the domain the reference's
    monotone-endpoint rules (collision_angles.cu:120,162,194,208,241) and case
    structure are not claimed, so points there are not checked."""
    vb = evaluator(family, cells)
    rng = np.random.default_rng(seed)
    bad, checked = [], 0
    for c, b in zip(cells, vb):
        if b & 7 != abi.V_PROVEN:
            continue
        for p in interior_points(c, family, 3 * k, rng):
            if not in_domain(family, p, lbrsp):
                continue
            holds, _, _ = model(family, p)
            checked += 1
            if not holds:
                bad.append((c.tolist(), p))
    return bad, checked


def wide_points(family, n, seed):
    """n points (la, r1, r2, r3), r1 >= r2 >= r3, of a box wider than the
    admissible domain (radii down to a fraction of the quarter lane), where the
    reusable angles and rho factors take both signs."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        la = rng.uniform(0.5, 0.99)
        w = 1 - la
        r1 = rng.uniform(0.12 * w, 0.5 * w)
        r2 = rng.uniform(0.08 * w, r1)
        r3 = rng.uniform(0.05 * w, r2) if abi.IS_3D[family] else 0.0
        out.append((la, r1, r2, r3))
    return np.array(out)


def term_mismatches(area_terms, family, points):
    """area_terms: (family, cells) -> ((n, 8, 2) pairs, (n,) counts), e.g. the
    oracle's area probe.  -> (mismatching points, points compared)."""
    pairs, counts = area_terms(family, point_cells(points))
    bad, compared = [], 0
    for p, got, k in zip(points, pairs, counts):
        if k < 0:
            continue  # decided before the criteria (fit check) or trapped: no area terms
        _, margin, want = model(family, p, physical=False)
        if margin < TOL:
            continue
        compared += 1
        if k != len(want):
            bad.append((p.tolist(), "checks", int(k), len(want)))
            continue
        for (lo_disk, hi_bound), (disk, bound) in zip(got[:k], want):
            ok = lo_disk <= disk and abs(disk - lo_disk) <= REL * abs(disk)
            ok = ok and hi_bound >= bound and abs(hi_bound - bound) <= REL * abs(bound) + 1e-15
            if not ok:
                bad.append((p.tolist(), float(lo_disk), float(disk), float(hi_bound), float(bound)))
                break
    return bad, compared
