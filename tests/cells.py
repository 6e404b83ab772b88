"""Seeded cell sets for the parity tests."""
import numpy as np

from circlepacking_amd import abi


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
