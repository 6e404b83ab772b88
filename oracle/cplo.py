"""ctypes binding of the CPU oracle (liboracle_cpl.so) — TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module,
as the checker; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

from circlepacking_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_cpl.so")
COUNT_LIB_PATH = os.path.join(HERE, "liboracle_cpl_count.so")
_libs = {}
_P_D = ctypes.POINTER(ctypes.c_double)
_P_U8 = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib(counting=False):
    """The oracle library; counting=True loads the FP64-operation-counting build."""
    path = COUNT_LIB_PATH if counting else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.cplo_prim.argtypes = [ctypes.c_int, _P_D, _P_D, _P_D, ctypes.c_size_t]
        L.cplo_geom.argtypes = [ctypes.c_int, _P_D, _P_D, ctypes.POINTER(ctypes.c_int), ctypes.c_size_t]
        L.cplo_interval_op.argtypes = [ctypes.c_int, _P_D, _P_D, _P_D, ctypes.POINTER(ctypes.c_int), ctypes.c_size_t]
        L.cplo_eval_kernel_body.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, _P_U8]
        L.cplo_ranges.argtypes = [ctypes.c_int, _P_D, _P_D, ctypes.c_size_t]
        L.cplo_area_terms.argtypes = [ctypes.c_int, _P_D, ctypes.c_size_t, _P_D, ctypes.POINTER(ctypes.c_int)]
        L.cplo_eval_cells.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P_D), ctypes.c_size_t, _P_U8,
                                      ctypes.c_int]
        L.cplo_refine_run.argtypes = [ctypes.POINTER(abi.RefineCfg), ctypes.POINTER(abi.RefineStats), _P_U8,
                                      ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint64)]
        L.cplo_first_problematic.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.POINTER(ctypes.c_int32), ctypes.c_int, _P_D,
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int), _P_D,
                                             ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
        _libs[path] = L
    return _libs[path]


def _dptr(a):
    return a.ctypes.data_as(_P_D)


def prim(op, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty_like(a)
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float64)
    code = abi.prim_code(op)
    rc = lib().cplo_prim(code, _dptr(a), None if bb is None else _dptr(bb), _dptr(out), a.size)
    assert rc == 0
    return out


GEOM = {"inner_inner": 0, "outer_inner": 1, "inner_outer": 2, "outer_outer": 3,
        "semicircle_outer": 4, "semicircle_inner": 5, "compute": 6}


def geom(fn, recs):
    """recs: (n, 6) = la.lb, la.ub, ra.lb, ra.ub, rb.lb, rb.ub -> ((n, 2) enclosure, (n,) trap codes)."""
    recs = np.ascontiguousarray(recs, dtype=np.float64)
    n = recs.shape[0]
    out = np.empty((n, 2), dtype=np.float64)
    trap = np.zeros(n, dtype=np.int32)
    rc = lib().cplo_geom(GEOM[fn], _dptr(recs), _dptr(out), trap.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    assert rc == 0
    return out, trap


def eval_kernel_body(family, cells):
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    out = np.empty(cells.shape[0], dtype=np.uint8)
    assert lib().cplo_eval_kernel_body(family, _dptr(cells), cells.shape[0], out.ctypes.data_as(_P_U8)) == 0
    return out


def area_terms(family, cells, library=None):
    """-> ((n, 8, 2) area comparisons (disk-area lower bound, bound on c x used
    area) of the output_criticals evaluation, (n,) counts; -1: no criteria ran)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    n = cells.shape[0]
    out = np.zeros((n, 16))
    cnt = np.zeros(n, dtype=np.int32)
    L = library or lib()
    assert L.cplo_area_terms(family, _dptr(cells), n, _dptr(out), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    return out.reshape(n, 8, 2), cnt


def ranges(family, recs):
    """recs: (n, 6) = la, r1, r2 intervals -> (n, 6) = r1 range(la), r2 range(la, r1), r3 range(la, r1, r2)."""
    recs = np.ascontiguousarray(recs, dtype=np.float64)
    out = np.empty_like(recs)
    assert lib().cplo_ranges(family, _dptr(recs), _dptr(out), recs.shape[0]) == 0
    return out


def eval_cells(family, soa, clip=True, nthreads=0):
    soa = np.ascontiguousarray(soa, dtype=np.float64)
    n = soa.shape[1]
    out = np.empty(n, dtype=np.uint8)
    ptrs = (_P_D * 8)(*[_dptr(soa[k]) for k in range(8)])
    assert lib().cplo_eval_cells(family, int(clip), ptrs, n, out.ctypes.data_as(_P_U8), nthreads) == 0
    return out


def refine_run(cfg, want_flags=True, root_begin=0, root_end=0, root_stride=1, nthreads=0, counting=False):
    """Synthetic refinement on the CPU over roots root_begin, +stride, ... < root_end.
    Returns (stats, flags) or, with counting=True, (stats, flags, flops[6, 8])."""
    st = abi.RefineStats()
    flops = np.zeros(48, dtype=np.uint64)
    n_roots = cfg.n_roots()
    end = root_end or n_roots
    count = max(0, (end - root_begin + root_stride - 1) // root_stride)
    flags = None
    fp = None
    if want_flags:
        flags = np.zeros((count, len(cfg.families())), dtype=np.uint8)
        fp = flags.ctypes.data_as(_P_U8)
    rc = lib(counting).cplo_refine_run(ctypes.byref(cfg), ctypes.byref(st), fp, root_begin, root_end, root_stride,
                                       nthreads, flops.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    assert rc == 0
    if counting:
        return st, flags, flops.reshape(6, 8)
    return st, flags


def first_problematic(family, outer_la, subdiv, nthreads=0):
    """-> (rc, first, evaluated, trap_code); every cell of the grid is evaluated."""
    r = mode_r(family, outer_la, subdiv, nthreads=nthreads)
    return r["rc"], r["first"], r["evaluated"], r["trap_code"]


def mode_r(family, outer_la, subdiv, output_criticals=False, rec_cap=1 << 16, nthreads=0):
    """One reference-grid launch of the oracle as a dict: rc, first, evaluated,
    trap_code, nofit_flags and, with output_criticals, records (k, 9) =
    kind, la, r1, r2, r3 of each printed line (sorted) and n_records."""
    first = ctypes.c_double()
    ev = ctypes.c_uint64()
    trap = ctypes.c_int()
    nrec = ctypes.c_uint64()
    nof = ctypes.c_uint64()
    sd = (ctypes.c_int32 * 4)(*subdiv)
    recs = np.zeros((rec_cap if output_criticals else 0, 9))
    rc = lib().cplo_first_problematic(family, outer_la[0], outer_la[1], sd, int(bool(output_criticals)),
                                      ctypes.byref(first), ctypes.byref(ev), ctypes.byref(trap),
                                      _dptr(recs) if output_criticals else None, recs.shape[0], ctypes.byref(nrec),
                                      ctypes.byref(nof), nthreads)
    out = {"rc": rc, "first": first.value, "evaluated": ev.value, "trap_code": trap.value,
           "nofit_flags": nof.value, "n_records": nrec.value}
    if output_criticals:
        out["records"] = recs[:min(nrec.value, rec_cap)]
    return out


def interval_op(op, a, b=None):
    """Oracle counterpart of circlepacking_amd.interval_op."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 2)
    n = a.shape[0]
    bb = None if b is None else np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float64)
    out = np.zeros((n, 2), dtype=np.float64)
    trap = np.zeros(n, dtype=np.int32)
    rc = lib().cplo_interval_op(abi.IV_OPS[op], _dptr(a), None if bb is None else _dptr(bb), _dptr(out),
                                trap.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    assert rc == 0
    return out, trap
