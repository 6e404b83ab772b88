"""Seeded operand generators shared by the primitive parity tests."""
import numpy as np

SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 3.0, 1e-300, -1e-300, 5e-324, -5e-324, 2.2250738585072014e-308,
                    1.7976931348623157e308, -1.7976931348623157e308, np.inf, -np.inf, 0.1, 1 / 3, 2 / 3,
                    0.9999999999999999, 1.0000000000000002, 1e16, 1e-16, 4.9e-324 * 3, 2.0 ** -1022 * 1.5])


def random_bits(rng, n):
    """Uniform random bit patterns without NaN/inf."""
    b = rng.integers(0, 2 ** 63, size=n, dtype=np.uint64)
    b = b % np.uint64(0x7FF0000000000000)
    sign = rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)
    return (b | sign).view(np.float64)


def moderate(rng, n):
    """Magnitudes in [2^-60, 2^60], random signs: the prover's operating range and beyond."""
    e = rng.uniform(-60, 60, size=n)
    s = rng.choice([-1.0, 1.0], size=n)
    return s * np.exp2(e)


def lane_like(rng, n):
    """Values shaped like the prover's operands (radii, lambda, angles, ratios)."""
    return rng.uniform(0.0, 2.0, size=n)


def short_mantissa(rng, n, bits=20):
    """Doubles with few mantissa bits: many exact products, quotients and sums."""
    m = rng.integers(1, 2 ** bits, size=n).astype(np.float64)
    e = rng.integers(-30, 30, size=n)
    s = rng.choice([-1.0, 1.0], size=n)
    return s * np.ldexp(m, e)


def binary_operands(seed, n):
    rng = np.random.default_rng(seed)
    k = n // 5
    a = np.concatenate([random_bits(rng, k), moderate(rng, k), lane_like(rng, k), short_mantissa(rng, k),
                        rng.choice(SPECIAL, size=n - 4 * k)])
    b = np.concatenate([random_bits(rng, k), moderate(rng, k), lane_like(rng, k), short_mantissa(rng, k),
                        moderate(rng, n - 4 * k)])
    # all pairs of specials
    sa, sb = np.meshgrid(SPECIAL, SPECIAL)
    a = np.concatenate([a, sa.ravel()])
    b = np.concatenate([b, sb.ravel()])
    # exact-quotient and exact-product constructions: a = q * b with short q
    q = short_mantissa(rng, k, bits=12)
    bb = short_mantissa(rng, k, bits=30)
    a = np.concatenate([a, q * bb])
    b = np.concatenate([b, bb])
    return a, b
