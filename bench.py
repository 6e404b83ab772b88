"""Benchmark of the lane-packing prover on one MI355X.

One step = one complete synthetic 4-parameter proof run: lambda in
[lambda_min, lambda_max], r1, r2 in [0.05, 0.5], r3 position in [0, 1] on a
64^4 root grid, families start / start&end 3-disk / middle / end 3-disk,
adaptive widest-axis bisection of undecided cells to depth 6; every root cell
is generated on the device, so the inputs are resident in HBM when timing
starts.  Prints ONE JSON line: value = (cell, family) leaves certified per
second of the whole run in the fastest form whose verdicts, counts, digests and
per-root flags are identical to the plain FP64 path.

Forms timed (each checked for identical results against the plain path):
  plain       one launch per family and round, host-synchronised sizes
  graph       all rounds captured once as a hipGraph with device-side sizes,
              hierarchical root generation, per-family parallel branches
  persistent  one persistent kernel per family draining a mixed-depth stack
  +fp32       the fastest batched form with the FP32 outward prefilter
  +learned    graph + FP32 prefilter gated by the learned proposal (MLP
              trained on the GPU from every record of the same 64^4 run,
              about 9.6e6, one per criteria evaluation)
  +learned order  graph with the two-check families' checks in the order the
              proposal predicts (same verdict kinds; failing-criterion details
              may differ), reported beside the identical forms
The learned split-axis policy changes the set of evaluated cells; it is timed
and reported separately (evaluations, unresolved leaves), not as `value`.

  python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-tune] [--no-learned] [--forms plain,...]
  python bench.py --traffic-probe   # 3 x 256 MiB device copy + one plain run (HBM counter calibration)
  python bench.py --graph-counters [graph|fp32|fp32+learned|learned-split] [--mlp FILE]
                                    # five proofs of one batched form (for rocprofv3); FILE from --save-mlp
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells certified/s; time-to-proof, tuned vs plain FP64 GPU, identical verdicts"
FP64_PEAK_TFLOPS = 78.6  # MI355X published FP64 peak (vector and matrix)
# Per-launch counter averages of the round kernels (rocprofv3 --pmc passes over
# `bench.py --traffic-probe`, summarised by `bench.py --summarize-pmc`).
COUNTERS_FILE = os.path.join(ROOT, "profiles", "k_round_counters_4param_64grid_depth6.json")
# VALU and FP64 instruction totals of one graph-form proof (rocprofv3 --pmc over
# `bench.py --graph-counters`, summed by `bench.py --summarize-graph-pmc`)
GRAPH_COUNTERS_FILE = os.path.join(ROOT, "profiles", "graph_counters_4param_64grid_depth6.json")
VALU_ISSUE_HZ = 1024 * 2.4e9  # 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD per 4 cycles at 2.4 GHz -> /4 below
FP64_LANE_OPS = 39.3e12       # 256 CUs x 64 FP64 lanes x 2.4 GHz (an fma counts once)


def kernel_groups(stats, cfg):
    """Launches grouped by kernel symbol: (family, root round?) -> [(round, seconds, verdict counts)]."""
    groups = {}
    for d in range(stats.rounds):
        for f in cfg.families():
            sec = stats.launch_seconds[d][f]
            if sec <= 0.0:
                continue
            counts = np.array(stats.family_verdicts[d][f][:8], dtype=np.float64)
            groups.setdefault((f, d == 0), []).append((d, sec, counts))
    return groups


def width_ratio_histogram(cpl, n=20000, seed=0):
    """FP32 / FP64 width ratio of the criteria's geometry intervals on boxes of
    the benchmark's domain (log2 bins)."""
    rng = np.random.default_rng(seed)
    la0 = rng.uniform(0.5, 0.98, n)
    law = rng.choice([0.0, 1e-5, 1e-3], n)
    lane = 1 - la0 - law
    ra0 = lane * rng.uniform(0.25, 0.45, n)
    rb0 = ra0 * rng.uniform(0.6, 1.0, n)
    w = lane * rng.choice([1e-6, 1e-4, 1e-3, 1e-2], n)
    recs = np.stack([la0, la0 + law, ra0, ra0 + w, rb0, rb0 + w * 0.5], 1)
    ratios = []
    for fn in ("inner_inner", "outer_inner", "inner_outer", "outer_outer", "semicircle_outer", "semicircle_inner"):
        g64, t64 = cpl.debug_geom(fn, recs, fp32=False)
        g32, t32 = cpl.debug_geom(fn, recs, fp32=True)
        ok = (t64 == 0) & (t32 == 0) & (g64[:, 1] > g64[:, 0])
        ratios.append((g32[ok, 1] - g32[ok, 0]) / (g64[ok, 1] - g64[ok, 0]))
    r = np.concatenate(ratios)
    edges = [1.0, 1.001, 1.01, 1.1, 2.0, 4.0, 16.0, np.inf]
    hist, _ = np.histogram(r, bins=edges)
    return {
        "samples": int(r.size),
        "bins": [f"[{a:g}, {b:g})" for a, b in zip(edges[:-1], edges[1:])],
        "counts": [int(c) for c in hist],
        "median": float(np.median(r)),
        "p99": float(np.percentile(r, 99)),
        "below_1": int(np.sum(r < 1.0)),
    }


def reference_grid_launches(cpl, abi):
    """The reference's own kernels: one launch per family over [lambda_min,
    lambda_max] on its full grids (cpl_first_problematic_ex, the call its prover
    makes per lambda interval), every cell evaluated (cells/s), then the same
    launch with the early exit the reference's threads take."""
    reference_grid = {}
    lam = (abi.LAMBDA_MIN, abi.LAMBDA_MAX)
    cpl.first_problematic_ex(0, lam, exhaustive=True)  # warm-up
    for f in range(6):
        t0 = time.perf_counter()
        first, info, _ = cpl.first_problematic_ex(f, lam, exhaustive=True)
        dt = time.perf_counter() - t0
        ev = info["cells_evaluated"]
        t0 = time.perf_counter()
        first_x, info_x, _ = cpl.first_problematic_ex(f, lam)
        dt_x = time.perf_counter() - t0
        reference_grid[abi.FAMILY_NAMES[f]] = {
            "ms": dt * 1e3, "cells": ev, "cells_per_s": ev / dt, "first_problematic_la": first,
            "early_exit": {"ms": dt_x * 1e3, "cells_evaluated": info_x["cells_evaluated"],
                           "rows_skipped": info_x["rows_skipped"], "same_first": first_x == first}}
    return reference_grid


def summarize_pmc(dirs):
    """Per-launch averages of every counter of every kernel in the rocprofv3
    counter CSVs under `dirs` (a counter is summed over its dimensions within a
    dispatch, then averaged over the dispatches of the kernel).  FETCH_SIZE and
    WRITE_SIZE are calibrated against the 256 MiB device copies of --traffic-probe
    (the `__amd_rocclr_copyBuffer` dispatches of the largest grid; the library's
    own small copies are left out), each of which reads and writes 256 MiB: the
    scales turn the KiB counts into bytes moved, and `hbm_bytes` is their scaled
    sum."""
    import csv
    import glob

    per_dispatch = {}  # (file, dispatch) -> {"name":..., counter: value}
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as fh:
                for row in csv.DictReader(fh):
                    key = (path, int(row["Dispatch_Id"]))
                    e = per_dispatch.setdefault(key, {"name": row["Kernel_Name"], "grid": int(row["Grid_Size"])})
                    e[row["Counter_Name"]] = e.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    copies = [e for e in per_dispatch.values() if e["name"].startswith("__amd_rocclr_copyBuffer")]
    big = max((e["grid"] for e in copies), default=0)
    sums, counts = {}, {}
    for e in per_dispatch.values():
        name = e["name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
        if name == "__amd_rocclr_copyBuffer" and e["grid"] == big:
            name = "__amd_rocclr_copyBuffer_256MiB"
        acc = sums.setdefault(name, {})
        cnt = counts.setdefault(name, {})
        for k, v in e.items():
            if k in ("name", "grid"):
                continue
            acc[k] = acc.get(k, 0.0) + v
            cnt[k] = cnt.get(k, 0) + 1
    out = {n: {k: acc[k] / counts[n][k] for k in acc} for n, acc in sums.items()}
    out = {n: dict(v, launches=max(counts[n].values())) for n, v in out.items()}
    copy = out.get("__amd_rocclr_copyBuffer_256MiB", {})
    copy_kib = 256 * 1024.0
    cal = {"copy_bytes": 256 << 20}
    if copy.get("FETCH_SIZE"):
        cal["fetch_scale"] = copy_kib / copy["FETCH_SIZE"]
    if copy.get("WRITE_SIZE"):
        cal["write_scale"] = copy_kib / copy["WRITE_SIZE"]
    for v in out.values():
        if "FETCH_SIZE" in v and "WRITE_SIZE" in v and "fetch_scale" in cal and "write_scale" in cal:
            v["hbm_bytes"] = 1024.0 * (v["FETCH_SIZE"] * cal["fetch_scale"] + v["WRITE_SIZE"] * cal["write_scale"])
    out["_calibration"] = cal
    return out


GRAPH_RUNS = 5  # proofs run by --graph-counters (each call replays the whole proof once)


def summarize_graph_pmc(dirs):
    """Instruction totals of one graph-form proof: every counter summed over
    all dispatches of each --pmc pass of `bench.py --graph-counters`, divided
    by GRAPH_RUNS (each refinement call, the two capturing ones included,
    executes the proof once; the library's small statistics kernels are
    included and negligible)."""
    import csv
    import glob

    tot = {}
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as fh:
                for row in csv.DictReader(fh):
                    if row["Kernel_Name"].startswith("__amd_rocclr"):
                        continue
                    tot[row["Counter_Name"]] = tot.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    out = {k: v / GRAPH_RUNS for k, v in tot.items()}
    out["fp64_insts"] = sum(out.get(k, 0.0) for k in ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
                                                      "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64"))
    out["runs"] = GRAPH_RUNS
    return out


def summarize_graph_pmc_by_kernel(dirs):
    """Per-kernel totals of the same passes (template arguments dropped: the
    instances of one kernel are summed), per proof."""
    import csv
    import glob

    tot = {}
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as fh:
                for row in csv.DictReader(fh):
                    name = row["Kernel_Name"].split("(")[0].split("<")[0].strip()
                    if name.startswith("void "):
                        name = name[len("void "):]
                    if name.startswith("__amd_rocclr"):
                        continue
                    k = tot.setdefault(name, {})
                    k[row["Counter_Name"]] = k.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    return {n: {c: v / GRAPH_RUNS for c, v in k.items()} for n, k in sorted(tot.items())}


# batched forms that `bench.py --graph-counters FORM` runs (options beside the
# graph form's graph + hierarchical roots + no digests)
GRAPH_FORMS = ("graph", "fp32", "fp32+learned", "learned-split", "learned-order")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--variant", default="tuned", choices=["plain", "tuned"],
                    help="which form `value` reports; every form is timed and checked for identical results")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (256 threads, 1 cell/thread)")
    ap.add_argument("--graph-shape", default=None,
                    help="block,cells_per_thread of the graph form instead of the autotuner's (e.g. 64,2)")
    ap.add_argument("--no-learned", action="store_true", help="skip training and timing the learned proposal")
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--depth", type=int, default=6)
    ap.add_argument("--cpu-stride", type=int, default=2, help="CPU baseline runs every S-th root cell")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="plain,graph,persistent,fp32",
                    help="forms to time besides plain (graph, persistent, fp32); the learned forms follow --no-learned")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-reference-grid", action="store_true",
                    help="skip the one-launch-per-family timing on the reference grids")
    ap.add_argument("--reference-grid-only", action="store_true",
                    help="only the one-launch-per-family timing on the reference grids (for rocprofv3), then exit")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run a 256 MiB device copy three times and one plain run, then exit (for rocprofv3 --pmc)")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="DIR",
                    help="rocprofv3 --pmc output directories (csv) of --traffic-probe runs: write the per-launch "
                         "counter averages of every kernel as JSON to --out, then exit (no GPU needed)")
    ap.add_argument("--out", default=None, help="output file of --summarize-pmc / --summarize-graph-pmc")
    ap.add_argument("--graph-counters", nargs="?", const="graph", default=None, choices=GRAPH_FORMS,
                    help="run GRAPH_RUNS proofs of one batched form (--graph-shape, default 64,2; the learned forms "
                         "load --mlp), then exit (for rocprofv3)")
    ap.add_argument("--summarize-graph-pmc", nargs="+", metavar="DIR",
                    help="rocprofv3 --pmc output directories of --graph-counters runs: write the instruction totals "
                         "of one proof as JSON to --out, then exit (no GPU needed)")
    ap.add_argument("--by-kernel", action="store_true",
                    help="with --summarize-graph-pmc: the totals per kernel as well")
    ap.add_argument("--key", default=None,
                    help="with --summarize-graph-pmc: store the totals under this key of the --out file, keeping "
                         "the file's other keys")
    ap.add_argument("--save-mlp", default=None, metavar="FILE",
                    help="write the proposals the benchmark trains to FILE (.npz, one array per family)")
    ap.add_argument("--mlp", default=None, metavar="FILE", help="proposals for --graph-counters (from --save-mlp)")
    args = ap.parse_args()
    if args.summarize_graph_pmc:
        summary = summarize_graph_pmc(args.summarize_graph_pmc)
        if args.by_kernel:
            summary["by_kernel"] = summarize_graph_pmc_by_kernel(args.summarize_graph_pmc)
        if args.key:
            merged = {}
            if args.out and os.path.exists(args.out):
                with open(args.out) as fh:
                    merged = json.load(fh)
            merged[args.key] = summary
            summary = merged
        text = json.dumps(summary, indent=1, sort_keys=True)
        if args.out:
            with open(args.out, "w") as fh:
                fh.write(text + "\n")
        print(text)
        return
    if args.summarize_pmc:
        summary = summarize_pmc(args.summarize_pmc)
        text = json.dumps(summary, indent=1, sort_keys=True)
        if args.out:
            with open(args.out, "w") as fh:
                fh.write(text + "\n")
        print(text)
        return
    if args.gpus != 1:
        print(f"bench.py: this prover runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the proof over several GPUs is not built", file=sys.stderr)
        sys.exit(2)

    import circlepacking_amd as cpl
    from circlepacking_amd import abi, learned

    if args.traffic_probe:
        import torch

        a = torch.ones(64 << 20, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        cpl.refine_run(abi.with_options(abi.cfg_4param(n=args.grid, max_depth=args.depth), abi.OPT_NO_DIGEST),
                       want_flags=False)
        print(json.dumps({"traffic_probe": True, "copy_bytes": 4 * a.numel()}), flush=True)
        return
    if args.graph_counters:
        b, c = (int(x) for x in (args.graph_shape or "64,2").split(","))
        base = abi.cfg_4param(n=args.grid, max_depth=args.depth)
        opts = abi.OPT_GRAPH | abi.OPT_HIER_ROOTS | abi.OPT_NO_DIGEST
        extra = {"graph": 0, "fp32": abi.OPT_FP32, "fp32+learned": abi.OPT_FP32 | abi.OPT_LEARNED,
                 "learned-split": abi.OPT_LEARNED_SPLIT, "learned-order": abi.OPT_LEARNED_ORDER}[args.graph_counters]
        if extra & (abi.OPT_LEARNED | abi.OPT_LEARNED_SPLIT | abi.OPT_LEARNED_ORDER):
            if not args.mlp:
                sys.exit("bench.py: --graph-counters " + args.graph_counters + " needs --mlp FILE (from --save-mlp)")
            with np.load(args.mlp) as z:
                for f in base.families():
                    cpl.mlp_install(f, z[abi.FAMILY_NAMES[f]])
        cfg = abi.with_options(base, opts | extra, block=b, cpt=c,
                               threshold=-0.5 if extra & abi.OPT_LEARNED else None)
        for _ in range(GRAPH_RUNS):
            cpl.refine_run(cfg, want_flags=False)
        print(json.dumps({"graph_counters": args.graph_counters, "runs": GRAPH_RUNS, "block": b,
                          "cells_per_thread": c}), flush=True)
        return
    if args.reference_grid_only:
        print(json.dumps({"reference_grid_launch": reference_grid_launches(cpl, abi)}), flush=True)
        return
    want = set(args.forms.split(","))

    G, H, PER, F32, L, LS, LO, ND = (abi.OPT_GRAPH, abi.OPT_HIER_ROOTS, abi.OPT_PERSISTENT, abi.OPT_FP32,
                                     abi.OPT_LEARNED, abi.OPT_LEARNED_SPLIT, abi.OPT_LEARNED_ORDER, abi.OPT_NO_DIGEST)
    base = abi.cfg_4param(n=args.grid, max_depth=args.depth)
    plain_ref = cpl.refine_run(abi.with_options(base, 0))  # digests + flags of the plain path

    def timed(cfg):
        for _ in range(args.warmup):
            cpl.refine_run(cfg, want_flags=False)
        t0 = time.perf_counter()
        st = None
        for _ in range(args.steps):
            st, _ = cpl.refine_run(cfg, want_flags=False)  # synchronous: returns after the device finished
        t1 = time.perf_counter()
        return (t1 - t0) * 1e3 / args.steps, st

    def identical(cfg):
        s, f = cpl.refine_run(abi.with_options(cfg, cfg.options & ~ND))
        return s.as_dict() == plain_ref[0].as_dict() and bool(np.array_equal(f, plain_ref[1]))

    forms = {}
    plain_cfg = abi.with_options(base, ND)
    ms_plain, st_plain = timed(plain_cfg)
    forms["plain"] = {"ms": ms_plain, "identical": True}

    graph_cfg = abi.with_options(base, G | H | ND)
    persist_cfg = abi.with_options(base, PER | H | ND, block=64, cpt=1)
    tune_graph = tune_persist = None
    if args.graph_shape:
        b, c = (int(x) for x in args.graph_shape.split(","))
        graph_cfg = abi.with_options(graph_cfg, graph_cfg.options, block=b, cpt=c)
    elif not args.no_tune and "graph" in want:
        tune_graph = cpl.autotune(graph_cfg, reps=3)
    if not args.no_tune and "persistent" in want:
        tune_persist = cpl.autotune(persist_cfg, reps=3)
    for name, cfg in (("graph", graph_cfg), ("persistent", persist_cfg)):
        if name not in want:
            continue
        ms, _ = timed(cfg)
        forms[name] = {"ms": ms, "identical": identical(cfg), "block": cfg.block, "cells_per_thread": cfg.cells_per_thread}

    batched = [k for k in ("graph", "persistent") if k in forms]
    if batched and "fp32" in want:
        # FP32 prefilter on the faster batched form
        batched_name = min(batched, key=lambda k: forms[k]["ms"])
        batched_cfg = graph_cfg if batched_name == "graph" else persist_cfg
        f32_cfg = abi.with_options(batched_cfg, batched_cfg.options | F32)
        ms, st = timed(f32_cfg)
        criteria_evals = sum(st.verdicts[d][v] for d in range(st.rounds)
                             for v in (abi.V_PROVEN, abi.V_CRITICAL, abi.V_TRAP))
        forms[batched_name + "+fp32"] = {
            "ms": ms, "identical": identical(f32_cfg), "fp32_tried": st.fp32_tried_total(),
            "fp32_settled": st.fp32_settled_total(),
            "fraction_settled_in_fp32": st.fp32_settled_total() / max(1, criteria_evals),
        }

    learned_info = None
    if not args.no_learned and "graph" in forms:
        t0 = time.perf_counter()
        # every record of the benchmark's own 64^4 run (about 9.6e6, one per
        # criteria evaluation), 8 epochs of Adam on the GPU
        train_cfg = abi.cfg_4param(n=args.grid, max_depth=args.depth)
        models = learned.train(cpl, train_cfg, families=base.families(), max_records=10_000_000, epochs=8, batch=4096)
        learned_info = {"train_seconds": time.perf_counter() - t0,
                        "records": {abi.FAMILY_NAMES[f]: m[1]["records"] for f, m in models.items()},
                        "final_loss": {abi.FAMILY_NAMES[f]: (m[1].get("loss") or [None])[-1] for f, m in models.items()}}
        if args.save_mlp:
            np.savez(args.save_mlp, **{abi.FAMILY_NAMES[f]: m[0] for f, m in models.items() if m[0] is not None})
        lg_cfg = abi.with_options(graph_cfg, G | H | ND | F32 | L, threshold=-0.5)
        ms, st = timed(lg_cfg)
        forms["graph+fp32+learned"] = {
            "ms": ms, "identical": identical(lg_cfg), "fp32_tried": st.fp32_tried_total(),
            "fp32_settled": st.fp32_settled_total(), "gate_threshold_logit": -0.5,
        }
        # learned check order: verdict kinds, counts, item digests and root flags
        # are the plain path's; the failing-criterion details may differ
        lo_cfg = abi.with_options(graph_cfg, G | H | ND | LO)
        ms, st = timed(lo_cfg)
        lo_s, lo_f = cpl.refine_run(abi.with_options(lo_cfg, G | H | LO))
        dk, dp = lo_s.as_dict(), plain_ref[0].as_dict()
        kind_keys = ("rounds", "evaluations", "verdicts", "digest", "family_verdicts", "total_evaluations",
                     "certified", "unresolved")
        forms["graph+learned order"] = {
            "ms": ms, "identical": False,
            "identical_verdict_kinds": all(dk[k] == dp[k] for k in kind_keys) and bool(np.array_equal(lo_f, plain_ref[1])),
            "identical_failing_criteria": dk["verdict_digest"] == dp["verdict_digest"],
            "note": "two-check families run first the check the proposals of a wave predict to fail first",
        }
        ls_cfg = abi.with_options(graph_cfg, G | H | ND | LS)
        ms, st = timed(ls_cfg)

        def volume(stats, cfg):  # unresolved volume, fraction of the box summed over the families
            return sum(stats.unresolved_volume(cfg).values())

        # the widest-axis rule one and two levels deeper (plain and graph forms):
        # the depth it needs to leave no more unresolved volume than the learned
        # split leaves at the benchmark's depth, and the time that takes
        target = volume(st, base)
        deeper, same_volume = [], None
        for d in (args.depth + 1, args.depth + 2):
            deep = abi.cfg_4param(n=args.grid, max_depth=d)
            rows = {}
            for fname, dcfg in (("plain", abi.with_options(deep, ND)),
                                ("graph", abi.with_options(deep, G | H | ND, block=graph_cfg.block,
                                                           cpt=graph_cfg.cells_per_thread))):
                ms_d, st_d = timed(dcfg)
                rows[fname] = {"max_depth": d, "form": fname, "ms": ms_d, "evaluations": st_d.total_evaluations,
                               "unresolved_volume": volume(st_d, deep)}
            deeper += [rows["plain"], rows["graph"]]
            reached = rows["plain"]["unresolved_volume"] <= target
            same_volume = {"reached": reached, "max_depth": d, "plain_ms": rows["plain"]["ms"],
                           "graph_ms": rows["graph"]["ms"], "speedup_vs_plain": rows["plain"]["ms"] / ms,
                           "speedup_vs_graph": rows["graph"]["ms"] / ms}
            if reached:
                break
        learned_info["learned_split"] = {
            "ms": ms, "evaluations": st.total_evaluations, "certified": st.certified, "unresolved": st.unresolved,
            "unresolved_volume": target,
            "plain_evaluations": st_plain.total_evaluations, "plain_unresolved": st_plain.unresolved,
            "plain_unresolved_volume": volume(st_plain, base),
            "widest_axis_deeper": deeper,
            "widest_axis_to_same_volume": same_volume,
            "note": "changes the set of evaluated cells (another bisection order); not an identical-verdict form. "
                    "unresolved_volume: volume of the leaves left unresolved, as a fraction of the box, summed over "
                    "the families; widest_axis_to_same_volume: the first depth at which the widest-axis rule leaves "
                    "no more unresolved volume, and its time against the learned split's; with reached false, the "
                    "deepest level tried, whose speedups are lower bounds",
        }
        for f in range(6):
            cpl.mlp_install(f, None)

    # value: the fastest identical form
    ok = {k: v for k, v in forms.items() if v["identical"]}
    tuned_name = min(ok, key=lambda k: ok[k]["ms"])
    if args.variant == "plain":
        name = "plain"
    else:
        name = tuned_name
    ms_per_step = forms[name]["ms"]
    certified = st_plain.certified
    proven_by_criteria = sum(st_plain.verdicts[d][abi.V_PROVEN] for d in range(st_plain.rounds))
    criteria_evaluations = sum(st_plain.verdicts[d][v] for d in range(st_plain.rounds)
                               for v in (abi.V_PROVEN, abi.V_CRITICAL, abi.V_TRAP))
    value = certified / (ms_per_step * 1e-3)

    # ---- CPU oracle on a stride sample: baseline timing + per-evaluation flop counts ----
    cpu = None
    cpu_parity = None
    flops_per_eval = None
    if not args.no_cpu:
        from oracle import cplo

        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        c0 = time.perf_counter()
        ost, oflags = cplo.refine_run(base, want_flags=True, root_stride=args.cpu_stride, nthreads=threads)
        c1 = time.perf_counter()
        cpu = {
            "value": ost.certified / (c1 - c0),
            "unit": "cells/s",
            "cores": threads,
            "host_cores": os.cpu_count(),
            "kind": "port",
            "sample": f"one root cell in {args.cpu_stride} of the same {args.grid}^4 grid ({ost.total_evaluations} "
                      f"evaluations, {ost.certified} certified leaves, {c1 - c0:.1f} s wall)",
        }
        # the same sample is a parity check of the GPU's per-(root, family) outcome flags
        cpu_parity = {"roots_compared": int(oflags.shape[0]),
                      "identical": bool(np.array_equal(oflags, plain_ref[1][::args.cpu_stride]))}
        fst, _, flops = cplo.refine_run(base, want_flags=False, root_stride=max(args.cpu_stride, 97),
                                        nthreads=threads, counting=True)
        counts = np.zeros((6, 8))
        for d in range(fst.rounds):
            counts += np.array([[fst.family_verdicts[d][f][v] for v in range(8)] for f in range(6)], dtype=np.float64)
        with np.errstate(invalid="ignore", divide="ignore"):
            flops_per_eval = np.where(counts > 0, flops / np.maximum(counts, 1), 0.0)

    # ---- roofline of the dominant kernel of the plain path (largest total device time) ----
    groups = kernel_groups(st_plain, plain_cfg)
    # the published FP64 peak, checked by an fma-throughput probe on this device
    peak_measured = cpl.fp64_peak(262144)  # ~30 ms of fma chains
    roof = None
    if groups:
        key = max(groups, key=lambda k: sum(s for _, s, _ in groups[k]))
        launches = groups[key]
        avg_sec = sum(s for _, s, _ in launches) / len(launches)
        f, is_root = key
        kname = f"k_round<{f}, {'true' if is_root else 'false'}, 256, false, 3>"  # the plain form's instance
        achieved = None
        if flops_per_eval is not None:
            flops_per_launch = sum(float(np.dot(c, flops_per_eval[f])) for _, _, c in launches) / len(launches)
            achieved = flops_per_launch / avg_sec / 1e12
        items = sum(float(c.sum()) for _, _, c in launches) / len(launches)
        traffic = fp64_util = valu_util = None
        if os.path.exists(COUNTERS_FILE):
            with open(COUNTERS_FILE) as fh:
                c = json.load(fh).get(kname, {})
            traffic = c.get("hbm_bytes")
            fp64 = sum(c.get(k, 0.0) for k in ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
                                                "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64"))
            if fp64:
                # wave instructions x 64 lanes against the FP64 lane-op peak
                fp64_util = fp64 * 64 / (avg_sec * FP64_LANE_OPS)
            if c.get("SQ_INSTS_VALU"):
                # every wave64 VALU instruction occupies its SIMD's VALU for 4 cycles
                valu_util = c["SQ_INSTS_VALU"] * 4 / (avg_sec * VALU_ISSUE_HZ)
        roof = {
            "bound": "fp64_valu",
            "pipe": "FP64 VALU (directed-rounding interval arithmetic; published FP64 vector peak)",
            "kernel": kname,
            "family": abi.FAMILY_NAMES[f],
            "launches": len(launches),
            "avg_launch_ms": avg_sec * 1e3,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "peak_measured": peak_measured,
            "unit": "TFLOP/s",
            "frac": (achieved / FP64_PEAK_TFLOPS) if achieved is not None else None,
            "traffic": traffic,
            "traffic_gbps": (traffic / avg_sec / 1e9) if traffic is not None else None,
            "fp64_valu_util": fp64_util,
            "valu_issue_util": valu_util,
            "worklist_items_per_launch": items,
        }

    # ---- the whole proof in the reported form: algorithmic flops of every
    # evaluation over the time to proof (all kernels, overlapped families), and
    # the VALU issue of the same run from its counter summary ----
    roof_proof = None
    if flops_per_eval is not None:
        total = np.zeros((6, 8))
        for d in range(st_plain.rounds):
            total += np.array([[st_plain.family_verdicts[d][f][v] for v in range(8)] for f in range(6)],
                              dtype=np.float64)
        proof_flops = float(np.sum(total * flops_per_eval))
        roof_proof = {"form": name, "flops": proof_flops, "time_ms": ms_per_step,
                      "achieved": proof_flops / (ms_per_step * 1e-3) / 1e12, "unit": "TFLOP/s",
                      "frac": proof_flops / (ms_per_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        if name == "graph" and os.path.exists(GRAPH_COUNTERS_FILE):
            with open(GRAPH_COUNTERS_FILE) as fh:
                gc = json.load(fh)
            roof_proof["valu_issue_util"] = gc["SQ_INSTS_VALU"] * 4 / (ms_per_step * 1e-3 * VALU_ISSUE_HZ)
            roof_proof["fp64_valu_util"] = gc["fp64_insts"] * 64 / (ms_per_step * 1e-3 * FP64_LANE_OPS)
            roof_proof["counters"] = os.path.relpath(GRAPH_COUNTERS_FILE, ROOT)

    reference_grid = None if args.no_reference_grid else reference_grid_launches(cpl, abi)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"4-parameter proof: lambda x r1 x r2 x r3-position box on a {args.grid}^4 root grid, "
                        f"families start/start_end_3d/middle/end_3d, widest-axis bisection to depth {args.depth}",
            "grid": [args.grid] * 4,
            "max_depth": args.depth,
            "variant": args.variant,
            "form": name,
        },
        "forms": forms,
        "tuned_form": tuned_name,
        "tuned_speedup": forms["plain"]["ms"] / forms[tuned_name]["ms"],
        # every form that keeps the plain path's verdicts does (the learned check
        # order is checked on verdict kinds instead)
        "identical_verdicts": all(v["identical"] if "identical_verdict_kinds" not in v else v["identical_verdict_kinds"]
                                  for v in forms.values()),
        "autotune_graph": tune_graph,
        "autotune_persistent": tune_persist,
        "learned": learned_info,
        "fp32_width_ratio": width_ratio_histogram(cpl) if "fp32" in want else None,
        "time_to_proof_ms": ms_per_step,
        "evaluations": st_plain.total_evaluations,
        "evaluations_per_s": st_plain.total_evaluations / (ms_per_step * 1e-3),
        "certified": certified,
        "certified_note": "leaves settled PROVEN, EMPTY or OOD (out-of-domain clipping included)",
        "proven_by_criteria": proven_by_criteria,
        "proven_by_criteria_per_s": proven_by_criteria / (ms_per_step * 1e-3),
        "criteria_evaluations": criteria_evaluations,
        "criteria_evaluations_per_s": criteria_evaluations / (ms_per_step * 1e-3),
        "unresolved": st_plain.unresolved,
        "roofline": roof,
        "roofline_proof": roof_proof,
        "cpu_baseline": cpu,
        "cpu_parity": cpu_parity,
        "reference_grid_launch": reference_grid,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
