"""bench.py's command line on a machine without a GPU: a run on more than one
GPU is refused with a message, before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_learned_counter_run_needs_proposals():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--graph-counters", "learned-split"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "needs --mlp" in r.stderr


def test_graph_pmc_summaries(tmp_path):
    """Counter sums of --graph-counters passes: per proof, rocclr copies left
    out, template instances of one kernel summed in the per-kernel table."""
    sys.path.insert(0, ROOT)
    import bench

    d = tmp_path / "pass" / "host"
    d.mkdir(parents=True)
    rows = [("k_round<0, false, 256, false, 4>(cpl::RefineParams)", "SQ_INSTS_VALU", 100.0),
            ("k_round<3, false, 256, false, 4>(cpl::RefineParams)", "SQ_INSTS_VALU", 50.0),
            ("k_round<3, false, 256, false, 4>(cpl::RefineParams)", "SQ_INSTS_VALU_FMA_F64", 20.0),
            ("k_propose<3>(cpl::RefineParams)", "SQ_INSTS_VALU", 10.0),
            ("__amd_rocclr_copyBuffer", "SQ_INSTS_VALU", 1000.0)]
    with open(d / "x_counter_collection.csv", "w") as fh:
        fh.write("Kernel_Name,Counter_Name,Counter_Value\n")
        for k, c, v in rows:
            fh.write(f'"{k}",{c},{v}\n')
    tot = bench.summarize_graph_pmc([str(tmp_path / "pass")])
    n = bench.GRAPH_RUNS
    assert tot["SQ_INSTS_VALU"] == 160.0 / n
    assert tot["fp64_insts"] == 20.0 / n
    per = bench.summarize_graph_pmc_by_kernel([str(tmp_path / "pass")])
    assert set(per) == {"k_round", "k_propose"}
    assert per["k_round"]["SQ_INSTS_VALU"] == 150.0 / n
