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
        L.cpl_eval_cells.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P_D), ctypes.c_size_t, _P_U8]
        L.cpl_eval_kernel_body.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, _P_U8]
        L.cpl_refine_run.argtypes = [ctypes.POINTER(abi.RefineCfg), ctypes.POINTER(abi.RefineStats), _P_U8]
        L.cpl_autotune.argtypes = [ctypes.POINTER(abi.RefineCfg), ctypes.c_int, ctypes.POINTER(abi.TunePoint),
                                   ctypes.POINTER(ctypes.c_int)]
        L.cpl_first_problematic.argtypes = [ctypes.c_int, abi.Iv, ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
                                            _P_D, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_size_t]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CplError(rc, lib().cpl_last_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(_P_D)


def prim(op, a, b=None):
    """Elementwise primitive on the GPU (op name from abi.PRIM_OPS)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float64)
    _check(lib().cpl_debug_prim(abi.PRIM_OPS[op], _dptr(a), None if bb is None else _dptr(bb), _dptr(out), a.size))
    return out


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


def autotune(cfg, reps=3):
    """Pick block size and cells-per-thread for cfg (in place); returns the timing table."""
    pts = (abi.TunePoint * 20)()
    n = ctypes.c_int()
    _check(lib().cpl_autotune(ctypes.byref(cfg), reps, pts, ctypes.byref(n)))
    return [(p.block, p.cells_per_thread, p.seconds) for p in pts[: n.value]]
