"""ctypes mirror of include/cpl.h (structures, enums and configuration presets)."""
import ctypes

START, START_END_2D, START_END_3D, MIDDLE, END_2D, END_3D = range(6)
FAMILY_NAMES = ["start", "start_end_2d", "start_end_3d", "middle", "end_2d", "end_3d"]
IS_3D = {START: True, START_END_2D: False, START_END_3D: True, MIDDLE: True, END_2D: False, END_3D: True}

V_PROVEN, V_CRITICAL, V_NOFIT, V_EMPTY, V_OOD, V_TRAP = range(6)
VERDICT_NAMES = ["PROVEN", "CRITICAL", "NOFIT", "EMPTY", "OOD", "TRAP"]

RF_CRITICAL, RF_NOFIT, RF_TRAP, RF_EMPTY, RF_OOD, RF_PROVEN, RF_NOFIT_FLAG = 1, 2, 4, 8, 16, 32, 64

CPL_OK = 0
ERR_NAMES = {-1: "HIP", -2: "ARG", -3: "TRAP", -4: "NOFIT", -5: "CAPACITY", -6: "NO_DEVICE"}
MAX_ROUNDS = 16

OPT_GRAPH, OPT_PERSISTENT, OPT_HIER_ROOTS, OPT_FP32, OPT_LEARNED, OPT_LEARNED_SPLIT, OPT_NO_DIGEST = 1, 2, 4, 8, 16, 32, 256
OPT_LEARNED_ORDER = 64

MLP_INPUTS, MLP_TARGETS, MLP_NPARAM = 8, 8, 5256
MLP_H, MLP_OUT = 64, 8
# parameter offsets: W1[8][64], b1[64], W2[64][64], b2[64], W3[64][8], b3[8]
MLP_W1, MLP_B1 = 0, 512
MLP_W2, MLP_B2 = 576, 4672
MLP_W3, MLP_B3 = 4736, 5248

LAMBDA_MIN = 0.5
LAMBDA_MAX = 0.9900000000000001

# reference subdivision counts per family: la, r1, r2, r3
REF_SUBDIV = {
    START: (256, 256, 256, 64),
    START_END_2D: (4096, 4096, 4096, 1),
    START_END_3D: (256, 256, 256, 64),
    MIDDLE: (512, 256, 256, 64),
    END_2D: (4096, 1024, 1024, 1),
    END_3D: (1024, 256, 128, 128),
}

PRIM_OPS = {
    "add_rd": 0, "add_ru": 1, "sub_rd": 2, "sub_ru": 3, "mul_rd": 4, "mul_ru": 5,
    "div_rd": 6, "div_ru": 7, "sqrt_rd": 8, "sqrt_ru": 9, "asin_rn": 10,
    "dasin_rd": 11, "dasin_ru": 12, "lbrsp": 13, "div_rn": 14, "sqrt_rn": 15,
}
PRIM_TRIG_OPS = {
    "sin_rn": 29, "cos_rn": 30, "acos_rn": 31, "dsin_rd": 32, "dsin_ru": 33, "dcos_rd": 34, "dcos_ru": 35,
    "dacos_rd": 36, "dacos_ru": 37,
}
# Atomic_ordered_double's transform and its inverse (the u64 image carried in a double's bits)
PRIM_ORDERED_OPS = {"ordered": 38, "ordered_inv": 39, "rcp_hw": 40, "rsq_hw": 41}
IV_OPS = {
    "add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "square": 5, "abs": 6, "reciprocal": 7,
    "div_scalar": 8, "scalar_div": 9, "sin": 10, "cos": 11, "asin": 12, "acos": 13, "mul_scalar": 14, "neg": 15,
}
IV_BINARY = {"add", "sub", "mul", "div", "div_scalar", "scalar_div", "mul_scalar"}


def prim_code(op):
    for table in (PRIM_OPS, PRIM32_OPS, PRIM_TRIG_OPS, PRIM_ORDERED_OPS):
        if op in table:
            return table[op]
    raise KeyError(op)


# FP32 prefilter primitives (operands and results are float values carried in doubles)
PRIM32_OPS = {
    "add_rd32": 16, "add_ru32": 17, "sub_rd32": 18, "sub_ru32": 19, "mul_rd32": 20, "mul_ru32": 21,
    "div_rd32": 22, "div_ru32": 23, "asin32_rn": 24, "dasin32_rd": 25, "dasin32_ru": 26,
    "f32_down": 27, "f32_up": 28,
}


class Iv(ctypes.Structure):
    _fields_ = [("lb", ctypes.c_double), ("ub", ctypes.c_double)]


MR_OUTPUT_CRITICALS, MR_EXHAUSTIVE = 1, 2
MR_MAX_RECORDS = 1 << 20
NOALLOC_MAX_LA, NOALLOC_MAX_RECORDS = 4096, 4096


class ModeRInfo(ctypes.Structure):
    _fields_ = [("cells_evaluated", ctypes.c_uint64), ("rows_skipped", ctypes.c_uint64),
                ("nofit_flags", ctypes.c_uint64), ("lines_found", ctypes.c_uint64),
                ("lines_written", ctypes.c_uint64), ("text_bytes_needed", ctypes.c_uint64),
                ("trap_code", ctypes.c_int32), ("fit_check_failed", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefineCfg(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_int32 * 4),
        ("box_lb", ctypes.c_double * 4),
        ("box_ub", ctypes.c_double * 4),
        ("family_mask", ctypes.c_uint32),
        ("clip", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("split_mode", ctypes.c_int32),
        ("split_axes", ctypes.c_uint32),
        ("options", ctypes.c_uint32),
        ("block", ctypes.c_int32),
        ("cells_per_thread", ctypes.c_int32),
        ("prefilter_depths", ctypes.c_uint32),
        ("learned_threshold", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 2),
    ]

    def families(self):
        return [f for f in range(6) if self.family_mask & (1 << f)]

    def n_roots(self):
        n = 1
        for g in self.grid:
            n *= g
        return n


class MlpTrainCfg(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("weight_decay", ctypes.c_float), ("reserved", ctypes.c_int32)]


class TunePoint(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("cells_per_thread", ctypes.c_int32), ("seconds", ctypes.c_double)]


class RefineStats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("evaluations", ctypes.c_uint64 * MAX_ROUNDS),
        ("verdicts", (ctypes.c_uint64 * 8) * MAX_ROUNDS),
        ("digest", ctypes.c_uint64 * MAX_ROUNDS),
        ("verdict_digest", ctypes.c_uint64 * MAX_ROUNDS),
        ("family_verdicts", ((ctypes.c_uint64 * 8) * 6) * MAX_ROUNDS),
        ("launch_seconds", (ctypes.c_double * 6) * MAX_ROUNDS),
        ("total_evaluations", ctypes.c_uint64),
        ("certified", ctypes.c_uint64),
        ("unresolved", ctypes.c_uint64),
        ("seconds", ctypes.c_double),
        ("eval_seconds", ctypes.c_double),
        ("peak_items", ctypes.c_uint64),
        ("fp32_settled", (ctypes.c_uint64 * 6) * MAX_ROUNDS),
        ("fp32_tried", (ctypes.c_uint64 * 6) * MAX_ROUNDS),
    ]

    def fp32_settled_total(self):
        return sum(self.fp32_settled[d][f] for d in range(MAX_ROUNDS) for f in range(6))

    def fp32_tried_total(self):
        return sum(self.fp32_tried[d][f] for d in range(MAX_ROUNDS) for f in range(6))

    def unresolved_volume(self, cfg):
        """Volume of the unresolved leaves as a fraction of the root box, per
        family of cfg: a leaf of depth d covers 2^-(d b) of its root cell (b
        axes bisected per split: 1 in split mode 0, the eligible axes in mode
        1); unresolved leaves are CRITICAL at the maximum depth, NOFIT and TRAP
        at any depth.  Independent of the split order, so runs that bisect
        different axes compare by it."""
        return self._leaf_volume(cfg, certified=False)

    def certified_volume(self, cfg):
        """Volume of the PROVEN, EMPTY and out-of-domain leaves, as unresolved_volume."""
        return self._leaf_volume(cfg, certified=True)

    def _leaf_volume(self, cfg, certified):
        bits = 1 if cfg.split_mode == 0 else bin(cfg.split_axes).count("1")
        out = {}
        for f in cfg.families():
            v = 0.0
            for d in range(self.rounds):
                fv = self.family_verdicts[d][f]
                if certified:
                    n = fv[V_PROVEN] + fv[V_EMPTY] + fv[V_OOD]
                else:
                    n = fv[V_NOFIT] + fv[V_TRAP] + (fv[V_CRITICAL] if d == cfg.max_depth else 0)
                v += n * 2.0 ** (-d * bits)
            out[FAMILY_NAMES[f]] = v / cfg.n_roots()
        return out

    def as_dict(self):
        r = self.rounds
        return {
            "rounds": r,
            "evaluations": list(self.evaluations[:r]),
            "verdicts": [list(self.verdicts[d][:6]) for d in range(r)],
            "digest": list(self.digest[:r]),
            "verdict_digest": list(self.verdict_digest[:r]),
            "family_verdicts": [[list(self.family_verdicts[d][f][:6]) for f in range(6)] for d in range(r)],
            "total_evaluations": self.total_evaluations,
            "certified": self.certified,
            "unresolved": self.unresolved,
        }


def make_cfg(grid, box_lb, box_ub, families, clip, max_depth, split_mode, split_axes, options=0, block=0, cpt=0):
    c = RefineCfg()
    c.grid[:] = list(grid)
    c.box_lb[:] = list(box_lb)
    c.box_ub[:] = list(box_ub)
    c.family_mask = sum(1 << f for f in families)
    c.clip = int(clip)
    c.max_depth = int(max_depth)
    c.split_mode = int(split_mode)
    c.split_axes = int(split_axes)
    c.options = int(options)
    c.block = int(block)
    c.cells_per_thread = int(cpt)
    return c


def with_options(cfg, options, block=None, cpt=None, prefilter_depths=None, threshold=None):
    """Copy of cfg with other execution options (and optionally shape)."""
    c = RefineCfg.from_buffer_copy(cfg)
    c.options = int(options)
    if threshold is not None:
        c.learned_threshold = float(threshold)
    if prefilter_depths is not None:
        c.prefilter_depths = int(prefilter_depths)
    if block is not None:
        c.block = int(block)
    if cpt is not None:
        c.cells_per_thread = int(cpt)
    return c


def cfg_small_2param(raw=False, n=256, max_depth=2):
    """2-parameter radius domain [0.05,0.5]^2 at lambda = 0.5, n x n grid, 2x2 splits."""
    return make_cfg((1, n, n, 1), (LAMBDA_MIN, 0.05, 0.05, 0.0), (LAMBDA_MIN, 0.5, 0.5, 1.0),
                    (START_END_2D, END_2D), not raw, max_depth, 1, 0b0110)


def cfg_4param(n=64, max_depth=6, families=(START, START_END_3D, MIDDLE, END_3D), grid=None):
    """4-parameter box: lambda in [lambda_min, lambda_max], r1, r2 in [0.05, 0.5],
    r3 position in [0, 1]; widest-axis bisection of undecided cells."""
    g = grid or (n, n, n, n)
    return make_cfg(g, (LAMBDA_MIN, 0.05, 0.05, 0.0), (LAMBDA_MAX, 0.5, 0.5, 1.0),
                    families, True, max_depth, 0, 0b1111)
