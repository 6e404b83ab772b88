"""The MLP oracle's gradients against central finite differences, and its
loss on closed-form cases (pins the checker of the GPU training kernels)."""
import numpy as np

from oracle import mlp_ref


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 8))
    t = np.zeros((n, 8))
    t[:, 0] = rng.integers(0, 2, n)
    t[:, 1:5] = rng.choice([0.0, 0.5, 1.0], (n, 4))
    t[:, 5] = rng.integers(0, 2, n)
    # verdict bytes: PROVEN, CRITICAL at criterion 1 or 2, out of domain
    t[:, 6] = rng.choice([0, 1 | 16, 1 | 32, 4], n)
    w = rng.normal(0, 0.4, mlp_ref.NPARAM)
    return x, t, w


def test_gradient_matches_finite_differences():
    x, t, w = _data(1, 40)
    _, g = mlp_ref.loss_and_grad(x, t, w)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(mlp_ref.NPARAM, 150, replace=False), [0, 511, 512, 575, 576, 4671, 4672,
                                                                           4735, 4736, 5247, 5248, 5255]])
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (mlp_ref.loss_and_grad(x, t, wp)[0] - mlp_ref.loss_and_grad(x, t, wm)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_loss_closed_forms():
    x, t, _ = _data(3, 25)
    w = np.zeros(mlp_ref.NPARAM)
    loss, g = mlp_ref.loss_and_grad(x, t, w)
    # zero network: y = 0, both BCE terms = log 2, squared error = sum t_a^2 / 4
    want = np.mean(2 * np.log(2.0) + t[:, 5] * np.sum(t[:, 1:5] ** 2, axis=1) / 4)
    assert abs(loss - want) < 1e-12
    # bias gradient of the gate output: mean(sigmoid(0) - t0)
    assert abs(g[5248] - np.mean(0.5 - t[:, 0])) < 1e-12
    # bias gradient of the check-order output: mean(sigmoid(0) - s), s = 1 only
    # for CRITICAL at criterion 2 (verdict byte 1 | 2 << 4)
    s = (t[:, 6] == (1 | 32)).astype(np.float64)
    assert abs(g[5248 + 5] - np.mean(0.5 - s)) < 1e-12
    assert 0 < s.sum() < len(s)


def test_pack_roundtrip_and_layout():
    w = np.arange(mlp_ref.NPARAM, dtype=np.float64)
    p = mlp_ref.unpack(w)
    assert p["W1"][1, 0] == 64 and p["b1"][0] == 512 and p["W2"][0, 1] == 577 and p["W3"][0, 1] == 4737
    assert np.array_equal(mlp_ref.pack(p), w)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.0000001, 0.0], dtype=np.float32)
    got = mlp_ref.bf16(v)
    assert got[0] == 1.0 and got[1] == 1.0  # tie to even
    assert got[2] == 1.0 + 2 ** -6 or got[2] == 1.0 + 2 ** -7
    assert got[3] == -3.0 and got[4] == 0.0
