"""Learned proposal: collect training records from a refinement run, train the
per-family MLP on the GPU, and install it for CPL_OPT_LEARNED runs.

No counterpart in the reference.  The proposal only decides whether the FP32
prefilter runs before the FP64 criteria on a cell (output 0), scores split
axes (outputs 1..4, used by CPL_OPT_LEARNED_SPLIT) and predicts that the second
check of a two-check family fails first (output 5, used by
CPL_OPT_LEARNED_ORDER); every verdict comes from the interval evaluation.
"""
import numpy as np

from . import abi


def init_params(seed=0):
    """He-normal hidden layers, small output layer, zero biases."""
    rng = np.random.default_rng(seed)
    w = np.zeros(abi.MLP_NPARAM, dtype=np.float32)
    w[abi.MLP_W1:abi.MLP_B1] = rng.normal(0, np.sqrt(2.0 / 8), 8 * 64)
    w[abi.MLP_W2:abi.MLP_B2] = rng.normal(0, np.sqrt(2.0 / 64), 64 * 64)
    w[abi.MLP_W3:abi.MLP_B3] = rng.normal(0, 0.5 * np.sqrt(1.0 / 64), 64 * 8)
    return w


def family_cfg(cfg, family):
    c = abi.RefineCfg.from_buffer_copy(cfg)
    c.family_mask = 1 << family
    c.options = 0
    return c


def train_family(cpl, cfg, family, max_records=2_000_000, keep_every=1, epochs=8, batch=4096, lr=2e-3, seed=0):
    """Records of `family` under cfg, shuffled, trained from init_params(seed).
    Returns (params, info)."""
    feat, targ = cpl.mlp_collect(family_cfg(cfg, family), max_records, keep_every)
    info = {"records": int(feat.shape[0])}
    if feat.shape[0] < 64:
        return None, info
    perm = np.random.default_rng(seed + 1).permutation(feat.shape[0])
    feat, targ = feat[perm], targ[perm]
    params, losses = cpl.mlp_train(feat, targ, init_params(seed), epochs=epochs, batch=batch, lr=lr)
    info["loss"] = [float(v) for v in losses]
    info["positive_fraction"] = float(targ[:, 0].mean())
    return params, info


def gate_quality(cpl, params, feat, targ, threshold=0.0):
    """Precision / recall of the gate (BF16 MFMA inference) against the FP32 outcome."""
    y = cpl.mlp_infer(feat, params)[:, 0]
    pred = y > threshold
    pos = targ[:, 0] > 0.5
    tp = float(np.sum(pred & pos))
    return {"precision": tp / max(1.0, float(pred.sum())), "recall": tp / max(1.0, float(pos.sum())),
            "predicted_fraction": float(pred.mean())}


def train(cpl, cfg, families=None, install=True, **kw):
    """Trains one proposal per family of cfg; returns {family: (params, info)}."""
    fams = families if families is not None else cfg.families()
    out = {}
    for f in fams:
        params, info = train_family(cpl, cfg, f, **kw)
        out[f] = (params, info)
        if install:
            cpl.mlp_install(f, params if params is not None else init_params())
    return out
