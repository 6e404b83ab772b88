"""The C++ lambda-queue driver (run_kernel_until_no_criticals semantics of
run_kernel.cuh:61-132) driven by a synthetic launcher on the CPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def trace_bin(tmp_path_factory):
    out = tmp_path_factory.mktemp("drv") / "driver_trace"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(out), os.path.join(ROOT, "tests", "driver_trace.cpp")],
                   check=True)
    return str(out)


def run(trace_bin, interrupt_at=-1):
    p = subprocess.run([trace_bin, str(interrupt_at)], capture_output=True, text=True, check=True)
    lines = p.stdout.strip().splitlines()
    launches = [tuple(float(x) for x in l.split()) for l in lines[:-1]]
    return launches, int(lines[-1].split()[1]), p.stderr


def test_queue_certifies_the_whole_lambda_range(trace_bin):
    launches, ok, log = run(trace_bin)
    assert ok == 1
    assert all(out == 0 for *_, out in launches)
    # processed in order of lower bound (min-heap), within the outer interval
    lbs = [l for l, *_ in launches]
    assert lbs == sorted(lbs) and lbs[0] == 0.5 and max(u for _, u, _, _ in launches) == 0.9900000000000001
    # certified prefixes [lb, first) of successive launches tile [0.5, 0.99] without gaps
    covered = []
    for l, u, f, _ in launches:
        covered.append((l, u if f == float("inf") else f))
    covered.sort()
    assert covered[0][0] == 0.5
    reach = 0.5
    for l, u in covered:
        assert l <= reach
        reach = max(reach, u)
    assert reach == 0.9900000000000001
    # every failing launch re-queues exactly per the split rule
    assert log.count("Handling lambda interval") == len(launches)


def test_split_rule(trace_bin):
    launches, _, _ = run(trace_bin)
    queued = {(l, u) for l, u, _, _ in launches[1:]}
    for l, u, f, _ in launches:
        if f == float("inf"):
            continue
        if (u - l) >= 2 * (u - f):
            assert (f, u) in queued
        else:
            mid = 0.5 * (f + u)
            assert (f, mid) in queued and (mid, u) in queued


def test_sigint_reruns_first_remaining_interval_with_output(trace_bin):
    full, _, _ = run(trace_bin)
    k = 3
    launches, ok, log = run(trace_bin, interrupt_at=k)
    assert ok == 0
    assert "We were interrupted" in log
    # launches 0..k as in the uninterrupted run, then one re-run with output
    assert launches[:k + 1] == full[:k + 1]
    last = launches[-1]
    assert last[3] == 1
    interrupted = launches[k]
    if interrupted[2] != float("inf"):
        assert last[:2] == interrupted[:2]


@pytest.fixture(scope="module")
def ckpt_bin(tmp_path_factory):
    out = tmp_path_factory.mktemp("drv") / "driver_checkpoint"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(out), os.path.join(ROOT, "tests", "driver_checkpoint.cpp")],
                   check=True)
    return str(out)


def _launches(stdout):
    return [l for l in stdout.strip().splitlines() if not l.startswith("result")]


@pytest.mark.parametrize("stop", [1, 7, 15])
def test_checkpoint_resume_reproduces_the_uninterrupted_trace(ckpt_bin, tmp_path, stop):
    full = subprocess.run([ckpt_bin, "run", str(tmp_path / "full.q"), "-1"], capture_output=True, text=True, check=True)
    path = str(tmp_path / "part.q")
    a = subprocess.run([ckpt_bin, "run", path, str(stop)], capture_output=True, text=True)
    assert a.returncode == 3
    b = subprocess.run([ckpt_bin, "resume", path], capture_output=True, text=True, check=True)
    assert _launches(a.stdout) + _launches(b.stdout) == _launches(full.stdout)
    assert b.stdout.strip().splitlines()[-1] == "result 1"
    # the finished run leaves an empty queue; resuming it launches nothing
    c = subprocess.run([ckpt_bin, "resume", path], capture_output=True, text=True, check=True)
    assert c.stdout.strip() == "result 1"


def test_interval_text_roundtrip_is_exact(ckpt_bin):
    import numpy as np

    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.uniform(-1, 1, 2000), rng.standard_normal(500) * 1e300, [5e-324, -0.0, 0.1, 1 / 3]])
    vals = vals[np.isfinite(vals)]
    lo, hi = np.minimum(vals, vals[::-1]), np.maximum(vals, vals[::-1])
    text = "\n".join(f"[{float(a)!r}, {float(b)!r}]" for a, b in zip(lo, hi)) + "\n[0.25; 0.5]\n0.75\n"
    p = subprocess.run([ckpt_bin, "roundtrip"], input=text, capture_output=True, text=True, check=True)
    out = p.stdout.strip().splitlines()
    assert len(out) == len(lo) + 2
    for (a, b), line in zip(zip(lo, hi), out):
        x, y = line.strip()[1:-1].split(",")
        assert float(x) == a and float(y) == b
    assert out[-2].replace(" ", "") == "[0.25,0.5]" and out[-1].replace(" ", "") == "[0.75,0.75]"
