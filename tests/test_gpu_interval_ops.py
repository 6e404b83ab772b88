"""The interval surface on the GPU (cpl_interval_op and the trig primitives)
bit-exact against the oracle, traps included."""
import numpy as np
import pytest

from circlepacking_amd import abi

pytestmark = pytest.mark.gpu


def _intervals(seed, n):
    rng = np.random.default_rng(seed)
    a = np.concatenate([rng.uniform(-10, 10, n), rng.uniform(-1, 1, n), rng.uniform(-1e4, 1e4, n // 4),
                        np.ldexp(rng.uniform(-1, 1, n // 4), rng.integers(-1000, 1000, n // 4))])
    w = np.abs(a) * rng.choice([0.0, 1e-12, 1e-6, 1e-2, 0.5, 3.0], a.size)
    return np.stack([a, a + w], 1)


@pytest.mark.parametrize("op", sorted(abi.IV_OPS))
def test_interval_ops_bit_exact(cpl, cplo, op):
    a = _intervals(abi.IV_OPS[op] + 1, 40_000)
    b = _intervals(abi.IV_OPS[op] + 101, 40_000)[: len(a)] if op in abi.IV_BINARY else None
    if op in ("div_scalar", "mul_scalar"):
        b = np.repeat(b[:, :1], 2, axis=1)
    g, gt = cpl.interval_op(op, a, b)
    o, ot = cplo.interval_op(op, a, b)
    assert np.array_equal(gt, ot), op
    ok = gt == 0
    same = (g.view(np.uint64) == o.view(np.uint64)) | (np.isnan(g) & np.isnan(o))
    assert np.all(same[ok]), (op, np.argwhere(~same[ok])[:3])


@pytest.mark.parametrize("op", sorted(abi.PRIM_TRIG_OPS))
def test_trig_primitives_bit_exact(cpl, cplo, op):
    rng = np.random.default_rng(abi.PRIM_TRIG_OPS[op])
    if "acos" in op:
        x = np.concatenate([rng.uniform(-1, 1, 500_000), [-1.0, 1.0, 0.5, -0.5, 0.0]])
    else:
        x = np.concatenate([rng.uniform(0, 8, 500_000), np.linspace(0, 8, 10001)])
    assert np.array_equal(cpl.prim(op, x).view(np.uint64), cplo.prim(op, x).view(np.uint64)), op
