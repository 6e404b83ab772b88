/* cpl.h — C ABI of the MI355X lane-packing prover (libcpl_hip.so).
 *
 * The reference (phillip-keldenich/circlepacking, code/) exposes this path as
 * C++ symbols inside one CUDA executable: kernels of the form
 *     __global__ void kernel(IV outer_la, Atomic_ordered_double* first, bool output_criticals)
 * launched by run_kernel_until_no_criticals (lane_packing/run_kernel.cuh:61-132)
 * for six families (start_configuration.cu, middle_configuration.cu,
 * end_configuration.cu).  This header is the drop-in boundary that replaces
 * those launches: plain C types, caller-owned host buffers, 0 on success and a
 * negative CPL_ERR_* code on failure (no exceptions cross it; the C++ driver
 * re-throws std::system_error as the reference's throw_if_cuda_error does,
 * algcuda/cuda-cpp-wrapper/include/algcuda/error.hpp:84-98).
 *
 * Threading: one library context per process and device; calls are
 * serialised on one HIP stream owned by the library; not reentrant (the
 * reference driver keeps global SIGINT state and is not reentrant either,
 * run_kernel.cuh:38-39).
 */
#ifndef CPL_H_INCLUDED
#define CPL_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interval of doubles; layout-identical to algcuda::Interval<double>
 * (algcuda/cuda-utils/cpp/include/algcuda/interval.hpp:84-242, two doubles lb, ub). */
typedef struct {
  double lb, ub;
} cpl_iv;

/* The six lane-packing kernels of the reference, one per configuration family. */
typedef enum {
  CPL_START = 0,        /* start_configuration.cu:114-167  kernel_find_critical_intervals_start_configuration */
  CPL_START_END_2D = 1, /* start_configuration.cu:190-215  ..._start_end_configuration_2d */
  CPL_START_END_3D = 2, /* start_configuration.cu:267-298  ..._start_end_configuration_3d */
  CPL_MIDDLE = 3,       /* middle_configuration.cu:156-205 ..._middle_configuration */
  CPL_END_2D = 4,       /* end_configuration.cu:103-146    ..._end_configuration_2d */
  CPL_END_3D = 5,       /* end_configuration.cu:212-268    ..._end_configuration_3d */
  CPL_NUM_FAMILIES = 6
} cpl_family;

/* Per-cell verdict byte: bits 0-2 = verdict, bit 3 = "possibly does not fit"
 * flag of the start family (the reference prints and keeps evaluating,
 * start_configuration.cu:131-135), bits 4-7 = detail (trap code for
 * CPL_V_TRAP, 1-based failing criterion for CPL_V_CRITICAL). */
enum {
  CPL_V_PROVEN = 0,   /* every criterion of the family holds on the cell */
  CPL_V_CRITICAL = 1, /* some criterion fails: the reference's store_min(la.lb) */
  CPL_V_NOFIT = 2,    /* middle/end fit check fires: the reference traps */
  CPL_V_EMPTY = 3,    /* tightening empties r1 or r2 (end_configuration.cu:232-234) */
  CPL_V_OOD = 4,      /* a radius lies outside the family's admissible range */
  CPL_V_TRAP = 5      /* a device consistency check of the reference would trap() */
};
/* Trap codes (detail nibble of CPL_V_TRAP). */
enum {
  CPL_TRAP_ASIN_DOMAIN = 1,       /* interval.cuh:617-621 */
  CPL_TRAP_OBTUSE_EMPTY_X = 2,    /* collision_angles.cu:58-62 */
  CPL_TRAP_OBTUSE_EMPTY_QUOT = 3, /* collision_angles.cu:71-75 */
  CPL_TRAP_ACUTE_EMPTY_X = 4,     /* collision_angles.cu:91-101 */
  CPL_TRAP_SQRT_DOMAIN = 5        /* interval.cuh:339-343 */
};

/* Error codes. */
enum {
  CPL_OK = 0,
  CPL_ERR_HIP = -1,      /* a HIP runtime call failed; see cpl_last_error() */
  CPL_ERR_ARG = -2,      /* invalid argument */
  CPL_ERR_TRAP = -3,     /* a cell hit a reference trap condition (replaces trap()) */
  CPL_ERR_NOFIT = -4,    /* middle/end fit check fired (the reference traps) */
  CPL_ERR_CAPACITY = -5, /* a worklist outgrew its device allocation */
  CPL_ERR_NO_DEVICE = -6
};

/* Grid axes of a synthetic refinement run. */
enum { CPL_AXIS_LA = 0, CPL_AXIS_R1 = 1, CPL_AXIS_R2 = 2, CPL_AXIS_P3 = 3, CPL_NUM_AXES = 4 };

#define CPL_MAX_ROUNDS 16

/* Synthetic refinement run (a worklist of interval cells, one kernel launch
 * per family and round).  Axis la holds the lambda interval, r1 and r2 the
 * radius intervals, p3 a position in [0,1] that selects r3 inside the
 * family's admissible r3 range (for a root cell it reproduces the reference's
 * nested get_subinterval rule, subintervals.hpp:31-42). */
typedef struct {
  int32_t grid[CPL_NUM_AXES]; /* root cells per axis */
  double box_lb[CPL_NUM_AXES], box_ub[CPL_NUM_AXES];
  uint32_t family_mask;       /* bit f: evaluate family f */
  int32_t clip;               /* 1: intersect radii with the admissible ranges (else raw box) */
  int32_t max_depth;          /* number of refinement rounds after the root round (< CPL_MAX_ROUNDS) */
  int32_t split_mode;         /* 0: bisect the widest eligible axis; 1: bisect all eligible axes */
  uint32_t split_axes;        /* bitmask of eligible axes */
  uint32_t options;           /* CPL_OPT_* (0: the plain FP64 path, one launch per family and round) */
  int32_t block;              /* threads per block of the evaluation kernels (0: 256) */
  int32_t cells_per_thread;   /* items a thread takes per scheduling step (0: 1) */
  uint32_t prefilter_depths;  /* with CPL_OPT_FP32: bit d enables the prefilter at depth d (0: every depth) */
  float learned_threshold;    /* with CPL_OPT_LEARNED: the prefilter runs on cells whose gate logit exceeds this */
  int32_t reserved[2];
} cpl_refine_cfg;

/* Execution options of cpl_refine_run.  Every option except
 * CPL_OPT_LEARNED_SPLIT and CPL_OPT_LEARNED_ORDER changes scheduling or the
 * precision of intermediate filtering only: verdicts, counts and root flags
 * are identical to the plain path.  CPL_OPT_LEARNED_SPLIT changes which cells
 * are evaluated (another bisection order); each verdict is still the interval
 * evaluation's.  CPL_OPT_LEARNED_ORDER changes the order of the two checks of
 * the middle, end 2-disk and end 3-disk families: verdict kinds, counts, item
 * digests and root flags stay the plain path's, and only the failing-criterion
 * detail of a cell on which both checks fail (and a trap that the other order
 * would not reach) can differ.  Neither has a counterpart in the reference. */
enum {
  CPL_OPT_GRAPH = 1,       /* all rounds enqueued with device-side sizes, captured once as a hipGraph, replayed */
  CPL_OPT_PERSISTENT = 2,  /* one persistent kernel drains a mixed-depth queue (LDS-staged batches) */
  CPL_OPT_HIER_ROOTS = 4,  /* decide admissibility once per (la, r1, r2) root triple */
  CPL_OPT_FP32 = 8,        /* FP32 outward-rounded prefilter; FP64 certifies what it leaves undecided */
  CPL_OPT_LEARNED = 16,    /* MLP proposal decides per cell whether the FP32 screen runs before the FP64 criteria */
  CPL_OPT_LEARNED_SPLIT = 32, /* mode-0 splits take the MLP's best-scored axis; changes the leaf set, not soundness */
  CPL_OPT_LEARNED_ORDER = 64, /* a wave evaluates first the check its cells' proposals predict to fail first */
  /* The learned options apply from depth 1 on (the root round runs without a
   * proposal) and run with the plain and CPL_OPT_GRAPH forms. */
  CPL_OPT_NO_DIGEST = 256  /* skip the per-item digests (timing runs) */
};

/* Per (root cell, family) outcome flags. */
enum {
  CPL_RF_CRITICAL = 1, /* a leaf stayed critical at max depth */
  CPL_RF_NOFIT = 2,
  CPL_RF_TRAP = 4,
  CPL_RF_EMPTY = 8,
  CPL_RF_OOD = 16,
  CPL_RF_PROVEN = 32,
  CPL_RF_NOFIT_FLAG = 64
};

typedef struct {
  int32_t rounds;                                  /* rounds executed */
  int32_t reserved0;
  uint64_t evaluations[CPL_MAX_ROUNDS];            /* (cell, family) items evaluated per round */
  uint64_t verdicts[CPL_MAX_ROUNDS][8];            /* per round, count of each verdict code */
  uint64_t digest[CPL_MAX_ROUNDS];                 /* sum of item hashes per round (order independent) */
  uint64_t verdict_digest[CPL_MAX_ROUNDS];         /* sum of hash(item, verdict byte) per round */
  uint64_t family_verdicts[CPL_MAX_ROUNDS][CPL_NUM_FAMILIES][8]; /* per round and family */
  double launch_seconds[CPL_MAX_ROUNDS][CPL_NUM_FAMILIES];        /* device time per launch (0 on the CPU) */
  uint64_t total_evaluations;
  uint64_t certified;                              /* leaves settled PROVEN, EMPTY or OOD */
  uint64_t unresolved;                             /* leaves CRITICAL at max depth, NOFIT or TRAP */
  double seconds;                                  /* device time of all rounds (HIP events) */
  double eval_seconds;                             /* device time of the criteria kernels only */
  uint64_t peak_items;                             /* largest worklist held on the device */
  uint64_t fp32_settled[CPL_MAX_ROUNDS][CPL_NUM_FAMILIES]; /* items the FP32 prefilter proved (CPL_OPT_FP32) */
  uint64_t fp32_tried[CPL_MAX_ROUNDS][CPL_NUM_FAMILIES];   /* items the FP32 prefilter ran on */
} cpl_refine_stats;

/* ---- lifecycle --------------------------------------------------------- */
/* Create the library context on the current HIP device (no counterpart in the
 * reference, which allocates per driver call: run_kernel.cuh:64). */
int cpl_init(void);
/* Release the context's device buffers, streams and events, the captured
 * graphs of the graph form and the installed proposals (install them again
 * after the next cpl_init); the no-allocation scratch stays (it is a device
 * array of the library). */
int cpl_shutdown(void);
/* Message of the last failing call (thread-local, never NULL). */
const char* cpl_last_error(void);

/* ---- the reference kernel contract --------------------------------------- */
/* One launch of family f over outer_la on the reference grid (subdivision
 * counts of start_configuration.cu:41-56, middle_configuration.cu:39-47,
 * end_configuration.cu:39-53), as run_kernel.cuh:81-84 does it: writes the
 * smallest la.lb of a cell failing a criterion, or DBL_MAX if none fails.
 * subdiv: NULL for the reference counts, else 4 counts la, r1, r2, r3 (r3
 * ignored by 2-d families) for truncated runs.  The same launch as
 * cpl_first_problematic_ex with flags = output_criticals ? CPL_MR_OUTPUT_CRITICALS
 * : 0 (see there: early exit, failure rule, printed lines); cells_evaluated_out
 * and out_text may be NULL.  Returns CPL_ERR_TRAP / CPL_ERR_NOFIT where the
 * reference would trap. */
int cpl_first_problematic(int family, cpl_iv outer_la, const int32_t* subdiv, int output_criticals,
                          double* first_problematic_out, uint64_t* cells_evaluated_out,
                          char* out_text, size_t out_text_cap);

/* Options and report of cpl_first_problematic_ex. */
enum {
  CPL_MR_OUTPUT_CRITICALS = 1, /* the reference's output_criticals = true (run_kernel.cuh:100-103) */
  CPL_MR_EXHAUSTIVE = 2        /* evaluate every cell (no early exit; cells_evaluated is the grid's count) */
};
#define CPL_MR_MAX_RECORDS (1u << 20) /* lines kept on the device per output-mode launch */
typedef struct {
  uint64_t cells_evaluated;   /* cells whose kernel body ran */
  uint64_t rows_skipped;      /* (la, r1) rows the early exit skipped */
  uint64_t nofit_flags;       /* start family: evaluated cells whose fit check fired (a printed line each) */
  uint64_t lines_found;       /* output mode: printed lines found (failing criteria + start no-fit lines) */
  uint64_t lines_written;     /* lines written to out_text (fewer than lines_found: truncated) */
  uint64_t text_bytes_needed; /* bytes the text of the lines kept on the device needs, with the final NUL */
  int32_t trap_code;          /* trap reported (CPL_TRAP_*; 0 none) */
  int32_t fit_check_failed;   /* the middle/end fit check was reported */
} cpl_mode_r_info;

/* One launch of family f over outer_la on the reference grid (the call of
 * run_kernel.cuh:81-84 and :100-103), with a report.
 *
 * Without CPL_MR_OUTPUT_CRITICALS the launch mirrors the reference kernels'
 * store_min + return (start_configuration.cu:142-145 and alike): every
 * reference thread walks its cells in increasing la and returns at its first
 * failing cell, so cells whose la.lb is at or above a failing la.lb found so
 * far cannot change the result and are skipped (rows_skipped; unless
 * CPL_MR_EXHAUSTIVE).  *first_problematic_out is the smallest la.lb of a cell
 * failing a criterion, DBL_MAX if none, either way.  A trap or a middle/end fit
 * failure is reported (CPL_ERR_TRAP, CPL_ERR_NOFIT) when its cell has la.lb
 * below that minimum: those cells are reached by the reference in every thread
 * mapping; above it the reference's outcome depends on its thread mapping.
 *
 * With CPL_MR_OUTPUT_CRITICALS every cell and every criterion is evaluated,
 * every trap and fit failure counts, and each line the reference would print
 * (one per failing criterion, and the start family's "possibly does not fit"
 * line) is appended to out_text in the reference's %.19g formats, in no
 * particular order.  At most CPL_MR_MAX_RECORDS lines are kept; lines_found
 * and lines_written show truncation, text_bytes_needed the buffer size the
 * kept lines need.  info may be NULL. */
int cpl_first_problematic_ex(int family, cpl_iv outer_la, const int32_t* subdiv, uint32_t flags,
                             double* first_problematic_out, cpl_mode_r_info* info, char* out_text,
                             size_t out_text_cap);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the launch of cpl_first_problematic_ex (same arguments,
 * flags, results and report) on a fixed device scratch array of the library
 * (CPL_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends) instead of
 * allocated buffers; it calls no hipMalloc, hipFree, memset or allocating
 * routine.  Limits: at most CPL_NOALLOC_MAX_LA lambda subdivisions (the
 * reference grids use at most 4096; more returns CPL_ERR_ARG), and in output
 * mode at most CPL_NOALLOC_MAX_RECORDS lines are kept (lines_found /
 * lines_written show truncation).  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU race (the library is not reentrant);
 * each call clears what it uses.  The library context (cpl_init) must belong to
 * the current device. */
#define CPL_NOALLOC_MAX_LA 4096u
#define CPL_NOALLOC_MAX_RECORDS 4096u
#define CPL_NOALLOC_SCRATCH_BYTES (CPL_NOALLOC_MAX_LA * 32u + 8448u + CPL_NOALLOC_MAX_RECORDS * 72u)
int cpl_first_problematic_noalloc(int family, cpl_iv outer_la, const int32_t* subdiv, uint32_t flags,
                                  double* first_problematic_out, cpl_mode_r_info* info, char* out_text,
                                  size_t out_text_cap);

/* ---- per-cell evaluation (criterion interface) -------------------------- */
/* Verdicts of family f for n cells given as 8 host arrays of n doubles
 * (la.lb, la.ub, r1.lb, r1.ub, r2.lb, r2.ub, p3.lb, p3.ub).  Equivalent to
 * evaluating the reference's handle_* criteria with the kernel-body gating. */
int cpl_eval_cells(int family, int clip, const double* const soa[8], size_t n, uint8_t* verdicts);

/* Reference kernel body on explicit cells: la, r1_, r2_ (grid subintervals
 * before tightening) and r3, 8 doubles per cell in that order (AoS). */
int cpl_eval_kernel_body(int family, const double* cells, size_t n, uint8_t* verdicts);

/* ---- synthetic refinement ---------------------------------------------- */
/* Runs the root round plus up to cfg->max_depth refinement rounds on the
 * device.  root_flags (may be NULL): n_roots * popcount(family_mask) bytes,
 * CPL_RF_* per (root, family) in increasing family order. */
int cpl_refine_run(const cpl_refine_cfg* cfg, cpl_refine_stats* stats, uint8_t* root_flags);

/* ---- autotuner ------------------------------------------------------------ */
typedef struct {
  int32_t block, cells_per_thread;
  double seconds;  /* median device time of the run with this shape */
} cpl_tune_point;

/* Times cfg (with its options) for every block size in {64,128,256,512,1024}
 * and cells-per-thread in {1,2,4,8} (median of `reps` runs each, after one
 * warm-up), writes the fastest shape into cfg->block / cfg->cells_per_thread,
 * and returns the table in `points` (up to 20 entries; may be NULL). */
int cpl_autotune(cpl_refine_cfg* cfg, int reps, cpl_tune_point* points, int* n_points);

/* ---- diagnostics (no reference counterpart; used by the parity tests) --- */
/* Elementwise primitive op on n host doubles: 0 add_rd 1 add_ru 2 sub_rd 3 sub_ru
 * 4 mul_rd 5 mul_ru 6 div_rd 7 div_ru 8 sqrt_rd 9 sqrt_ru 10 asin_rn
 * 11 dasin_rd 12 dasin_ru 13 lower_bound_radius_single_placement(a)
 * 14 div_rn 15 sqrt_rn. */
int cpl_debug_prim(int op, const double* a, const double* b, double* out, size_t n);

/* No counterpart in the reference: FP64 FMA throughput of the device (eight
 * independent fma chains per lane, `iters` fmas per chain), in TFLOP/s (fma =
 * 2 flops).  A check of the published FP64 peak that roofline figures use. */
int cpl_debug_fp64_peak(int iters, double* tflops);
/* FP32 prefilter primitives through the same entry point (operands and
 * results are float values carried in doubles): 16 add_rd 17 add_ru 18 sub_rd
 * 19 sub_ru 20 mul_rd 21 mul_ru 22 div_rd 23 div_ru 24 asin_rn 25 dasin_rd
 * 26 dasin_ru; 27 / 28: a double rounded down / up to float.  Round-to-nearest
 * and widened functions of the interval surface (argument in [0, 8] for sin and
 * cos): 29 sin_rn 30 cos_rn 31 acos_rn 32 dsin_rd 33 dsin_ru 34 dcos_rd
 * 35 dcos_ru 36 dacos_rd 37 dacos_ru.  The ordered minimum's bit map
 * (atomic_float.hpp:56-66): 38 the u64 image of a carried in out's bits,
 * 39 the double whose image is the u64 carried in a's bits.  The hardware
 * estimates the division and square root build on: 40 v_rcp_f64, 41 v_rsq_f64. */

/* Elementwise interval operations of the algcuda::Interval<double> surface
 * (interval.cuh) on n intervals; b is the second operand (binary ops; for the
 * scalar forms its lb is the scalar).  trap: n codes (CPL_TRAP_*; 0 none).
 * Differences to the reference, both in the enclosing direction: scalar x
 * interval rounds the upper bound up (interval.cuh:185-197 rounds it down) and
 * the reciprocal of a negative interval is [rd(1/ub), ru(1/lb)]
 * (interval.cuh:227-235 returns the inverted pair).  sin / cos / acos use the
 * library's own round-to-nearest functions widened by 3 ulps. */
enum {
  CPL_IV_ADD = 0, CPL_IV_SUB = 1, CPL_IV_MUL = 2, CPL_IV_DIV = 3,   /* interval.cuh:38-66, 171-182, 277-316 */
  CPL_IV_SQRT = 4, CPL_IV_SQUARE = 5, CPL_IV_ABS = 6,               /* :330-355, interval.hpp:216-232 */
  CPL_IV_RECIPROCAL = 7, CPL_IV_DIV_SCALAR = 8, CPL_IV_SCALAR_DIV = 9, /* :227-235, 261-275 */
  CPL_IV_SIN = 10, CPL_IV_COS = 11, CPL_IV_ASIN = 12, CPL_IV_ACOS = 13, /* :578-624 */
  CPL_IV_MUL_SCALAR = 14, CPL_IV_NEG = 15                             /* :185-197, interval.hpp:131 */
};
int cpl_interval_op(int op, const cpl_iv* a, const cpl_iv* b, cpl_iv* out, int* trap, size_t n);

/* Geometry of the criteria on n records of 6 doubles (la, ra, rb intervals):
 * fn 0 collision_angle_inner_inner 1 outer_inner 2 inner_outer (collision_angles.cu:119-205)
 * 3 outer_outer(ra, rb) (:207-231) 4 semicircle_angle_outer(ra) (:233-238)
 * 5 semicircle_angle_inner(la, ra) (:240-246) 6 compute_collision_angle(la, ra, rb) (:37-117);
 * 7 / 8 the same with only the lower / upper endpoint computed (the forms the
 * criteria call; the other endpoint of out is unspecified).
 * fp32 = 1 evaluates in the prefilter precision on the outward FP32 enclosure
 * of the inputs.  out: n x (lb, ub); trap: n trap codes (0: none). */
int cpl_debug_geom(int fn, int fp32, const double* in, double* out, int* trap, size_t n);

/* For n cells (8 doubles each, as cpl_eval_kernel_body): the FP32 criteria
 * result on the tightened cell (0: every criterion holds, k: criterion k fails
 * first, 16 + trap code; 255 when the cell is decided before the criteria)
 * and the FP64 verdict byte. */
int cpl_debug_prefilter(int family, const double* cells, size_t n, uint8_t* fp32_result, uint8_t* verdicts);

/* ---- learned proposal (no reference counterpart) ------------------------- */
/* An 8 -> 64 -> 64 -> 8 ReLU network per family.  Parameters, FP32, in this
 * order: W1[8][64], b1[64], W2[64][64], b2[64], W3[64][8], b3[8] (W[i][j]
 * connects input i to unit j).  Output 0: logit of "the FP32 prefilter proves
 * the cell"; outputs 1..4: predicted one-step settle score of splitting axis
 * la, r1, r2, p3. */
#define CPL_MLP_INPUTS 8
#define CPL_MLP_TARGETS 8
#define CPL_MLP_NPARAM 5256

/* Training records of a refinement run: every evaluated (cell, family) item
 * that reaches the criteria (PROVEN, CRITICAL or TRAP), one in keep_every by a
 * hash of the item.  features: n x 8 floats (per axis: midpoint mapped to
 * [-1, 1] over the box, log2(width / box extent) / 16 + 1 clamped to [-1, 1]);
 * targets:
 * n x 8 floats (0: FP32 prefilter proves the cell, 1..4: fraction of the two
 * children settled when splitting axis a (mode-0 runs, CRITICAL cells above
 * max depth), 5: 1 where 1..4 are defined, 6: FP64 verdict byte, 7: depth).
 * Runs the plain path; family_mask selects the families (records of one
 * family when one bit is set). */
int cpl_mlp_collect(const cpl_refine_cfg* cfg, uint64_t max_records, uint32_t keep_every, float* features,
                    float* targets, uint64_t* n_records);

typedef struct {
  int32_t epochs, batch;
  float lr, beta1, beta2, eps, weight_decay;
  int32_t reserved;
} cpl_mlp_train_cfg;

/* Trains params (in: initial, out: trained) with Adam on the GPU over n records
 * in the given order, minibatches of cfg->batch; epoch_loss (may be NULL):
 * cfg->epochs mean losses.  Loss: BCE on output 0 plus targets[5] * mean
 * squared error of outputs 1..4. */
int cpl_mlp_train(const float* features, const float* targets, size_t n, float* params,
                  const cpl_mlp_train_cfg* cfg, float* epoch_loss);
/* Mean loss, its gradient (CPL_MLP_NPARAM floats) and optionally the FP32
 * outputs (n x 8) of the training kernels over n records (gradient checks). */
int cpl_mlp_grad(const float* features, const float* targets, size_t n, const float* params, float* grad_out,
                 float* loss_out, float* outputs);
/* Outputs 0..5 (n x 6) of the in-kernel BF16 MFMA inference path. */
int cpl_mlp_infer(const float* features, size_t n, const float* params, float* out);
/* Installs (params != NULL) or removes the proposal of a family for the
 * CPL_OPT_LEARNED / CPL_OPT_LEARNED_SPLIT / CPL_OPT_LEARNED_ORDER runs. */
int cpl_mlp_install(int family, const float* params);

#ifdef __cplusplus
}
#endif
#endif
