"""MI355X-native lane-packing interval prover.

Host-side binding of libcpl_hip.so (include/cpl.h).  The HIP library is the
only compute path: if it is missing or no GPU is present, every call raises.
"""
import ctypes
import os

import numpy as np

from . import abi
from .abi import *  # noqa: F401,F403

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpl_hip.so")
_lib = None

_P_D = ctypes.POINTER(ctypes.c_double)
_P_U8 = ctypes.POINTER(ctypes.c_uint8)


class CplError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cpl error {code} ({abi.ERR_NAMES.get(code, '?')}): {msg}")
        self.code = code


def lib():
    """Load libcpl_hip.so (fails loudly if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() or python -m circlepacking_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.cpl_last_error.restype = ctypes.c_char_p
        L.cpl_debug_prim.argtypes = [ctypes.c_int, _P_D, _P_D, _P_D, ctypes.c_size_t]
        L.cpl_debug_fp64_peak.argtypes = [ctypes.c_int, _P_D]
        L.cpl_eval_cells.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P_D), ctypes.c_size_t, _P_U8]
        L.cpl_eval_kernel_body.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, _P_U8]
        L.cpl_debug_geom.argtypes = [ctypes.c_int, ctypes.c_int, _P_D, _P_D, ctypes.POINTER(ctypes.c_int),
                                     ctypes.c_size_t]
        L.cpl_debug_prefilter.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, _P_U8, _P_U8]
        L.cpl_interval_op.argtypes = [ctypes.c_int, _P_D, _P_D, _P_D, ctypes.POINTER(ctypes.c_int), ctypes.c_size_t]
        _P_F = ctypes.POINTER(ctypes.c_float)
        L.cpl_mlp_collect.argtypes = [ctypes.POINTER(abi.RefineCfg), ctypes.c_uint64, ctypes.c_uint32, _P_F, _P_F,
                                      ctypes.POINTER(ctypes.c_uint64)]
        L.cpl_mlp_train.argtypes = [_P_F, _P_F, ctypes.c_size_t, _P_F, ctypes.POINTER(abi.MlpTrainCfg), _P_F]
        L.cpl_mlp_grad.argtypes = [_P_F, _P_F, ctypes.c_size_t, _P_F, _P_F, _P_F, _P_F]
        L.cpl_mlp_infer.argtypes = [_P_F, ctypes.c_size_t, _P_F, _P_F]
        L.cpl_mlp_install.argtypes = [ctypes.c_int, _P_F]
        L.cpl_refine_run.argtypes = [ctypes.POINTER(abi.RefineCfg), ctypes.POINTER(abi.RefineStats), _P_U8]
        L.cpl_autotune.argtypes = [ctypes.POINTER(abi.RefineCfg), ctypes.c_int, ctypes.POINTER(abi.TunePoint),
                                   ctypes.POINTER(ctypes.c_int)]
        L.cpl_first_problematic.argtypes = [ctypes.c_int, abi.Iv, ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
                                            _P_D, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_size_t]
        L.cpl_first_problematic_ex.argtypes = [ctypes.c_int, abi.Iv, ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32,
                                               _P_D, ctypes.POINTER(abi.ModeRInfo), ctypes.c_char_p, ctypes.c_size_t]
        L.cpl_first_problematic_noalloc.argtypes = L.cpl_first_problematic_ex.argtypes
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CplError(rc, lib().cpl_last_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(_P_D)


def fp64_peak(iters=65536):
    """Measured FP64 FMA throughput of the device in TFLOP/s."""
    t = ctypes.c_double()
    _check(lib().cpl_debug_fp64_peak(int(iters), ctypes.byref(t)))
    return t.value


def prim(op, a, b=None):
    """Elementwise primitive on the GPU (op name from abi.PRIM_OPS)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float64)
    code = abi.prim_code(op)
    _check(lib().cpl_debug_prim(code, _dptr(a), None if bb is None else _dptr(bb), _dptr(out), a.size))
    return out


GEOM = {"inner_inner": 0, "outer_inner": 1, "inner_outer": 2, "outer_outer": 3,
        "semicircle_outer": 4, "semicircle_inner": 5, "compute": 6, "compute_lb": 7, "compute_ub": 8}


def debug_geom(fn, recs, fp32=False):
    """Geometry function `fn` (name from GEOM) on an (n, 6) array of interval
    records (la, ra, rb); returns ((n, 2) bounds, (n,) trap codes)."""
    recs = np.ascontiguousarray(recs, dtype=np.float64).reshape(-1, 6)
    n = recs.shape[0]
    out = np.zeros((n, 2), dtype=np.float64)
    trap = np.zeros(n, dtype=np.int32)
    _check(lib().cpl_debug_geom(GEOM[fn], int(bool(fp32)), _dptr(recs), _dptr(out),
                                trap.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n))
    return out, trap


def interval_op(op, a, b=None):
    """Elementwise interval operation (name from abi.IV_OPS) on (n, 2) arrays of
    [lb, ub]; returns ((n, 2) result, (n,) trap codes)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 2)
    n = a.shape[0]
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float64)
    out = np.zeros((n, 2), dtype=np.float64)
    trap = np.zeros(n, dtype=np.int32)
    _check(lib().cpl_interval_op(abi.IV_OPS[op], _dptr(a), None if bb is None else _dptr(bb), _dptr(out),
                                 trap.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n))
    return out, trap


def debug_prefilter(family, cells):
    """(FP32 criteria result, FP64 verdict byte) per cell of an (n, 8) array."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 8)
    n = cells.shape[0]
    r32 = np.zeros(n, dtype=np.uint8)
    v64 = np.zeros(n, dtype=np.uint8)
    _check(lib().cpl_debug_prefilter(int(family), _dptr(cells), n, r32.ctypes.data_as(_P_U8),
                                     v64.ctypes.data_as(_P_U8)))
    return r32, v64


def eval_cells(family, soa, clip=True):
    """Verdict bytes of `family` for cells given as an (8, n) float64 array."""
    soa = np.ascontiguousarray(soa, dtype=np.float64)
    assert soa.ndim == 2 and soa.shape[0] == 8
    n = soa.shape[1]
    out = np.empty(n, dtype=np.uint8)
    ptrs = (_P_D * 8)(*[_dptr(soa[k]) for k in range(8)])
    _check(lib().cpl_eval_cells(family, int(clip), ptrs, n, out.ctypes.data_as(_P_U8)))
    return out


def eval_kernel_body(family, cells):
    """Reference kernel body on explicit cells: (n, 8) array la, r1_, r2_, r3."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    assert cells.ndim == 2 and cells.shape[1] == 8
    out = np.empty(cells.shape[0], dtype=np.uint8)
    _check(lib().cpl_eval_kernel_body(family, _dptr(cells), cells.shape[0], out.ctypes.data_as(_P_U8)))
    return out


def refine_run(cfg, want_flags=True):
    """Synthetic refinement run; returns (RefineStats, root flags (n_roots, nfam) or None)."""
    st = abi.RefineStats()
    flags = None
    fp = None
    if want_flags:
        flags = np.zeros((cfg.n_roots(), len(cfg.families())), dtype=np.uint8)
        fp = flags.ctypes.data_as(_P_U8)
    _check(lib().cpl_refine_run(ctypes.byref(cfg), ctypes.byref(st), fp))
    return st, flags


def first_problematic(family, outer_la, subdiv=None, output_criticals=False, text_cap=1 << 20):
    """One reference-grid launch: (first problematic la.lb or DBL_MAX, cells evaluated, critical text).
    Raises CplError(-3/-4) where the reference would trap."""
    first = ctypes.c_double()
    ev = ctypes.c_uint64()
    sd = None
    if subdiv is not None:
        sd = (ctypes.c_int32 * 4)(*subdiv)
    buf = ctypes.create_string_buffer(text_cap) if output_criticals else None
    rc = lib().cpl_first_problematic(family, abi.Iv(*outer_la), sd, int(output_criticals), ctypes.byref(first),
                                     ctypes.byref(ev), buf, text_cap if buf is not None else 0)
    _check(rc)
    return first.value, ev.value, (buf.value.decode() if buf is not None else "")


def first_problematic_ex(family, outer_la, subdiv=None, output_criticals=False, exhaustive=False, text_cap=1 << 22,
                         noalloc=False):
    """One reference-grid launch with its report: (first problematic la.lb or
    DBL_MAX, info dict, printed lines as text).  Raises CplError(-3/-4) where
    the reference would trap; the error carries .first, .info and .text.
    noalloc: through cpl_first_problematic_noalloc (fixed device scratch)."""
    first = ctypes.c_double()
    info = abi.ModeRInfo()
    sd = None
    if subdiv is not None:
        sd = (ctypes.c_int32 * 4)(*subdiv)
    flags = (abi.MR_OUTPUT_CRITICALS if output_criticals else 0) | (abi.MR_EXHAUSTIVE if exhaustive else 0)
    buf = ctypes.create_string_buffer(text_cap) if output_criticals else None
    fn = lib().cpl_first_problematic_noalloc if noalloc else lib().cpl_first_problematic_ex
    rc = fn(family, abi.Iv(*outer_la), sd, flags, ctypes.byref(first), ctypes.byref(info), buf,
            text_cap if buf is not None else 0)
    text = buf.value.decode() if buf is not None else ""
    if rc != 0:
        err = CplError(rc, lib().cpl_last_error().decode(errors="replace"))
        err.info, err.first, err.text = info.as_dict(), first.value, text
        raise err
    return first.value, info.as_dict(), text


def autotune(cfg, reps=3):
    """Pick block size and cells-per-thread for cfg (in place); returns the timing table."""
    pts = (abi.TunePoint * 20)()
    n = ctypes.c_int()
    _check(lib().cpl_autotune(ctypes.byref(cfg), reps, pts, ctypes.byref(n)))
    return [(p.block, p.cells_per_thread, p.seconds) for p in pts[: n.value]]


# ---- learned proposal -------------------------------------------------------
def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, cols=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if cols is not None:
        a = a.reshape(-1, cols)
    return a


def mlp_collect(cfg, max_records, keep_every=1):
    """(features (n, 8), targets (n, 8)) float32 records of a plain run of cfg."""
    feat = np.zeros((int(max_records), 8), dtype=np.float32)
    targ = np.zeros((int(max_records), 8), dtype=np.float32)
    n = ctypes.c_uint64()
    _check(lib().cpl_mlp_collect(ctypes.byref(cfg), int(max_records), int(keep_every), _fptr(feat), _fptr(targ),
                                 ctypes.byref(n)))
    return feat[:n.value].copy(), targ[:n.value].copy()


def mlp_train(features, targets, params, epochs=5, batch=4096, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8,
              weight_decay=0.0):
    """Trains on the GPU; returns (trained params, per-epoch mean loss)."""
    x, t = _f32(features, 8), _f32(targets, 8)
    w = _f32(params).copy()
    tc = abi.MlpTrainCfg(epochs, batch, lr, beta1, beta2, eps, weight_decay, 0)
    losses = np.zeros(epochs, dtype=np.float32)
    _check(lib().cpl_mlp_train(_fptr(x), _fptr(t), x.shape[0], _fptr(w), ctypes.byref(tc), _fptr(losses)))
    return w, losses


def mlp_grad(features, targets, params, want_outputs=False):
    """(mean loss, gradient, outputs or None) of the GPU training kernels."""
    x, t, w = _f32(features, 8), _f32(targets, 8), _f32(params)
    g = np.zeros(abi.MLP_NPARAM, dtype=np.float32)
    loss = ctypes.c_float()
    out = np.zeros((x.shape[0], 8), dtype=np.float32) if want_outputs else None
    _check(lib().cpl_mlp_grad(_fptr(x), _fptr(t), x.shape[0], _fptr(w), _fptr(g), ctypes.byref(loss),
                              None if out is None else _fptr(out)))
    return float(loss.value), g, out


def mlp_infer(features, params):
    """Outputs 0..5 (n, 6) of the in-kernel BF16 MFMA inference."""
    x, w = _f32(features, 8), _f32(params)
    out = np.zeros((x.shape[0], 6), dtype=np.float32)
    _check(lib().cpl_mlp_infer(_fptr(x), x.shape[0], _fptr(w), _fptr(out)))
    return out


def mlp_install(family, params):
    """Installs (or with None removes) the proposal of one family."""
    if params is None:
        _check(lib().cpl_mlp_install(int(family), None))
    else:
        w = _f32(params)
        assert w.size == abi.MLP_NPARAM
        _check(lib().cpl_mlp_install(int(family), _fptr(w)))
