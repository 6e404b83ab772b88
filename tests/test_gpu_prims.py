"""GPU directed-rounding primitives, asin and the single-placement bound,
bit-exact against the CPU oracle (x86 MXCSR directed rounding) through the C ABI."""
import numpy as np
import pytest

from tests.operands import binary_operands, random_bits

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def assert_bit_exact(got, want, a, b=None, what=""):
    g, w = bits(got), bits(want)
    bad = np.nonzero((g != w) & ~(np.isnan(got) & np.isnan(want)))[0]
    if bad.size:
        i = bad[0]
        extra = "" if b is None else f", b={b[i]!r}"
        raise AssertionError(f"{what}: {bad.size} mismatches; first a={a[i]!r}{extra}: got {got[i]!r} want {want[i]!r}")


@pytest.mark.parametrize("op", ["add_rd", "add_ru", "sub_rd", "sub_ru", "mul_rd", "mul_ru", "div_rd", "div_ru"])
def test_binary_directed(cpl, cplo, op):
    a, b = binary_operands(seed=hash(op) % 2 ** 32, n=2_000_000)
    assert_bit_exact(cpl.prim(op, a, b), cplo.prim(op, a, b), a, b, op)


@pytest.mark.parametrize("op", ["sqrt_rd", "sqrt_ru"])
def test_sqrt_directed(cpl, cplo, op):
    rng = np.random.default_rng(7)
    a = np.abs(np.concatenate([random_bits(rng, 1_000_000), rng.uniform(0, 4, 1_000_000),
                               np.square(rng.integers(1, 2 ** 26, 100_000).astype(np.float64)),
                               [0.0, -0.0, 5e-324, 2.2250738585072014e-308, np.inf, 1.0, 4.0]]))
    a = np.concatenate([a, [-1.0, -5e-324]])
    assert_bit_exact(cpl.prim(op, a), cplo.prim(op, a), a, None, op)


def test_rn_division_and_sqrt_are_ieee(cpl):
    rng = np.random.default_rng(11)
    a, b = random_bits(rng, 1_000_000), random_bits(rng, 1_000_000)
    assert_bit_exact(cpl.prim("div_rn", a, b), a / b, a, b, "div_rn")
    x = np.abs(random_bits(rng, 1_000_000))
    assert_bit_exact(cpl.prim("sqrt_rn", x), np.sqrt(x), x, None, "sqrt_rn")


@pytest.mark.parametrize("op", ["asin_rn", "dasin_rd", "dasin_ru"])
def test_asin_bit_exact(cpl, cplo, op):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-1, 1, 1_000_000), rng.uniform(0.49, 0.51, 100_000), 1 - rng.uniform(0, 1e-6, 100_000),
                        rng.uniform(-1e-10, 1e-10, 10_000), [0.0, -0.0, 0.5, -0.5, 1.0, -1.0, 5e-324, 1.5, -2.0, np.nan]])
    assert_bit_exact(cpl.prim(op, x), cplo.prim(op, x), x, None, op)


def test_asin_endpoints(cpl):
    x = np.array([1.0, -1.0, 0.0, -0.0, 0.5])
    y = cpl.prim("asin_rn", x)
    assert y[0] == 1.5707963267948966 and y[1] == -1.5707963267948966 and y[2] == 0.0 and y[4] > 0.5
    u = np.linspace(-1, 1, 100001)
    assert not np.any(np.isnan(cpl.prim("asin_rn", u)))


def test_single_placement_bound(cpl, cplo):
    lam = np.concatenate([np.linspace(0.5, 0.9900000000000001, 2049), [0.0, 0.25, 0.999]])
    assert_bit_exact(cpl.prim("lbrsp", lam), cplo.prim("lbrsp", lam), lam, None, "lbrsp")


def test_fp64_peak_probe_is_plausible(cpl):
    # eight fma chains per lane on every CU: the published FP64 peak of an
    # MI355X is 78.6 TFLOP/s; the probe must land near it (clock and power
    # management move it by a few percent)
    t = cpl.fp64_peak(16384)
    assert 40.0 < t < 90.0, t
