"""Seeded cell sets for the parity tests."""
import numpy as np

from circlepacking_amd import abi
from tests import exact_ref


def c1_grid_soa(n=256, lam=abi.LAMBDA_MIN):
    """Root cells of the 2-parameter grid as an (8, n*n) SoA array (get_subinterval rule, fma form)."""
    edges = np.empty(n + 1)
    lo, hi = 0.05, 0.5
    diff = (hi - lo) / n
    # l_i = fma(diff, i, lo); numpy has no fma, so use exact rational rounding through float128-free path
    from fractions import Fraction
    for i in range(n + 1):
        edges[i] = float(Fraction(diff) * i + Fraction(lo)) if i < n else hi
    r1 = np.repeat(np.arange(n), n)
    r2 = np.tile(np.arange(n), n)
    soa = np.empty((8, n * n))
    soa[0] = lam
    soa[1] = lam
    soa[2] = edges[r1]
    soa[3] = edges[r1 + 1]
    soa[4] = edges[r2]
    soa[5] = edges[r2 + 1]
    soa[6] = 0.0
    soa[7] = 1.0
    return soa


def random_kernel_body_cells(seed, n, family):
    """Random (la, r1_, r2_, r3) boxes around the admissible band, plus degenerate and wide ones."""
    rng = np.random.default_rng(seed)
    la_lo = rng.uniform(0.5, 0.99, n)
    la_w = rng.choice([0.0, 1e-6, 1e-3, 1e-2], n)
    la = np.stack([la_lo, np.minimum(la_lo + la_w, 0.9900000000000001)], 1)
    lane = 1.0 - la_lo
    out = np.empty((n, 8))
    out[:, 0:2] = la
    for k, (flo, fhi) in enumerate([(0.2, 0.40), (0.15, 0.33), (0.1, 0.33)]):
        lo = lane * rng.uniform(flo, fhi, n)
        w = lane * rng.choice([0.0, 1e-4, 1e-2, 0.05, 0.3], n)
        out[:, 2 + 2 * k] = lo
        out[:, 3 + 2 * k] = lo + w
    # keep r1 >= r2 >= r3 roughly ordered as in the reference grids for most cells
    m = rng.uniform(size=n) < 0.8
    r = np.sort(out[m][:, [2, 4, 6]], axis=1)[:, ::-1]
    rw = out[m][:, [3, 5, 7]] - out[m][:, [2, 4, 6]]
    out[np.nonzero(m)[0], 2] = r[:, 0]
    out[np.nonzero(m)[0], 4] = r[:, 1]
    out[np.nonzero(m)[0], 6] = r[:, 2]
    out[np.nonzero(m)[0], 3] = r[:, 0] + rw[:, 0]
    out[np.nonzero(m)[0], 5] = r[:, 1] + rw[:, 1]
    out[np.nonzero(m)[0], 7] = r[:, 2] + rw[:, 2]
    return out


def grid_cells(ref, f, n, seed, subdiv):
    """n random cells (la, r1_, r2_, r3; before tightening) of the reference's
    nested grid with the given subdivision counts (subintervals.hpp:31-42 and
    the families' range functions), lambda over [lambda_min, lambda_max]."""
    rng = np.random.default_rng(seed)
    nla, n1, n2, n3 = subdiv
    out = []
    while len(out) < n:
        ia, i1, i2, i3 = (int(rng.integers(k)) for k in (nla, n1, n2, n3))
        la = exact_ref.get_subinterval(exact_ref.IV(exact_ref.LAMBDA_MIN, exact_ref.LAMBDA_MAX), ia, nla)
        r1_ = exact_ref.get_subinterval(ref.compute_r1_range(la), i1, n1)
        r2_ = exact_ref.get_subinterval(ref.compute_r2_range(f, la, r1_), i2, n2)
        r3 = exact_ref.IV(0.0, 0.0)
        if f in exact_ref.THREE_D:
            r3_range = ref.compute_r3_range(f, la, r1_, r2_)
            if f == exact_ref.END_3D and r3_range.empty():
                continue
            r3 = exact_ref.get_subinterval(r3_range, i3, n3)
        out.append([la.lb, la.ub, r1_.lb, r1_.ub, r2_.lb, r2_.ub, r3.lb, r3.ub])
    return np.array(out)
