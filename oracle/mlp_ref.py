"""CPU restatement of the learned proposal's MLP (8 -> 64 -> 64 -> 8, ReLU).

Test infrastructure only: the checker of the GPU training kernels (forward,
loss, gradients) and of the in-kernel BF16 MFMA inference.  The proposal has
no counterpart in the reference.  Float64 throughout; its gradients are
pinned by central finite differences in tests/test_oracle_mlp.py.

Parameter vector (5256 floats): W1[8][64], b1[64], W2[64][64], b2[64],
W3[64][8], b3[8]; W[i][j] connects input i to unit j.  Loss per record:
BCE(y0, t0) + t5 * sum_{a=1..4} (y_a - t_a)^2 / 4 + BCE(y5, s), averaged over
records, where s = 1 when the FP64 verdict byte t6 is CRITICAL (kind 1) with
failing criterion 2 (detail nibble 2), else 0.
"""
import numpy as np

SHAPES = [("W1", (8, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 8)), ("b3", (8,))]
NPARAM = 5256


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = w[k:k + n].reshape(shape)
        k += n
    assert k == NPARAM
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def forward(x, w):
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ p["W1"] + p["b1"], 0.0)
    h2 = np.maximum(h1 @ p["W2"] + p["b2"], 0.0)
    y = h2 @ p["W3"] + p["b3"]
    return h1, h2, y


def loss_and_grad(x, t, w):
    """Mean loss and its gradient (float64, flat like the parameter vector)."""
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = x.shape[0]
    h1, h2, y = forward(x, w)
    y0, t0, m = y[:, 0], t[:, 0], t[:, 5]
    bce = np.maximum(y0, 0) - y0 * t0 + np.log1p(np.exp(-np.abs(y0)))
    d = y[:, 1:5] - t[:, 1:5]
    vb = t[:, 6].astype(np.int64)
    s5 = (((vb & 7) == 1) & (((vb >> 4) & 15) == 2)).astype(np.float64)
    y5 = y[:, 5]
    bce5 = np.maximum(y5, 0) - y5 * s5 + np.log1p(np.exp(-np.abs(y5)))
    loss = float(np.mean(bce + m * np.sum(d * d, axis=1) / 4.0 + bce5))
    dy = np.zeros_like(y)
    dy[:, 0] = (1.0 / (1.0 + np.exp(-y0)) - t0) / n
    dy[:, 1:5] = m[:, None] * d / 2.0 / n
    dy[:, 5] = (1.0 / (1.0 + np.exp(-y5)) - s5) / n
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    dh2 = (dy @ p["W3"].T) * (h2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ p["W2"].T) * (h1 > 0)
    g["W1"] = x.T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, pack(g)


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(x, w):
    """Outputs 0..5 as the MFMA inference computes them: BF16 weights and
    activations, wide accumulation, FP32 biases."""
    p = unpack(np.asarray(w, dtype=np.float32))
    h1 = np.maximum(bf16(x) @ bf16(p["W1"]) + p["b1"], 0.0)
    h2 = np.maximum(bf16(h1) @ bf16(p["W2"]) + p["b2"], 0.0)
    y = bf16(h2) @ bf16(p["W3"]) + p["b3"]
    return y[:, :6]
