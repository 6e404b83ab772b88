"""Mode R: the cpl_prover executable (HIP launches) produces the same lambda
trace as the same driver fed by the CPU oracle, on truncated reference grids."""
import os
import subprocess

import pytest

from circlepacking_amd import abi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROVER = os.path.join(ROOT, "circlepacking_amd", "cpl_prover")


@pytest.fixture(scope="module")
def oracle_driver(tmp_path_factory, cplo):
    out = tmp_path_factory.mktemp("drv") / "driver_oracle"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(out), os.path.join(ROOT, "tests", "driver_oracle.cpp"),
                    cplo.LIB_PATH, "-Wl,-rpath," + os.path.dirname(cplo.LIB_PATH)], check=True)
    return str(out)


def handled(stderr):
    return [l for l in stderr.splitlines() if l.startswith("Handling lambda interval")]


CASES = [
    (abi.START, "6,6,6,3", "0.5,0.6"),
    (abi.START_END_2D, "12,12,12,1", "0.5,0.52"),
    (abi.START_END_3D, "6,6,6,3", "0.55,0.6"),
    (abi.MIDDLE, "6,6,6,3", "0.5,0.56"),
    (abi.END_2D, "12,12,12,1", "0.5,0.6"),
    (abi.END_3D, "6,6,6,3", "0.5,0.6"),
]


@pytest.mark.parametrize("family,subdiv,lam", CASES)
def test_lambda_trace_matches_oracle_driver(oracle_driver, family, subdiv, lam):
    assert os.path.exists(PROVER), "cpl_prover is not built"
    limit = "40"
    g = subprocess.run([PROVER, "--family", str(family), "--subdiv", subdiv, "--lambda", lam, "--max-launches", limit],
                       capture_output=True, text=True, timeout=120)
    o = subprocess.run([oracle_driver, str(family), subdiv, lam, limit], capture_output=True, text=True, timeout=120)
    assert g.returncode in (0, 3), g.stderr[-2000:]
    assert o.returncode == g.returncode, (o.stderr[-500:], g.stderr[-500:])
    hg, ho = handled(g.stderr), handled(o.stderr)
    assert len(hg) >= 1 and hg == ho


def test_prover_banners_in_reference_order():
    g = subprocess.run([PROVER, "--subdiv", "4,4,4,2", "--lambda", "0.5,0.51", "--max-launches", "30"],
                       capture_output=True, text=True, timeout=120)
    text = g.stdout + g.stderr
    order = ["Start configuration...", "Start & end configuration, 2 disks...", "Start & end configuration, 3 disks ...",
             "Middle configuration ...", "End configuration (2 disks) ...", "End configuration (3 disks) ..."]
    pos = [text.find(b) for b in order]
    if g.returncode == 0:
        assert all(p >= 0 for p in pos) and pos == sorted(pos)
    else:
        assert g.returncode == 3 and pos[0] >= 0


def test_checkpoint_resume_continues_the_trace(tmp_path):
    """A run stopped after 5 launches and resumed from its checkpoint handles
    the same lambda intervals, in the same order, as an uninterrupted run."""
    fam, subdiv, lam = abi.MIDDLE, "6,6,6,3", "0.5,0.56"
    full = subprocess.run([PROVER, "--family", str(fam), "--subdiv", subdiv, "--lambda", lam, "--max-launches", "30"],
                          capture_output=True, text=True, timeout=120)
    ck = str(tmp_path / "queue.txt")
    a = subprocess.run([PROVER, "--family", str(fam), "--subdiv", subdiv, "--lambda", lam, "--max-launches", "5",
                        "--checkpoint", ck], capture_output=True, text=True, timeout=120)
    ha = handled(a.stderr)
    if a.returncode == 0:  # the whole stage took at most 5 launches
        assert ha == handled(full.stderr)
        return
    assert a.returncode == 3 and len(ha) == 6  # the 6th interval was popped, then the limit stopped the launch
    b = subprocess.run([PROVER, "--family", str(fam), "--subdiv", subdiv, "--lambda", lam, "--max-launches", "25",
                        "--resume", ck, "--checkpoint", ck], capture_output=True, text=True, timeout=120)
    assert b.returncode == full.returncode
    hf = handled(full.stderr)
    assert ha[:5] + handled(b.stderr) == hf[:5 + len(handled(b.stderr))]
    assert len(ha[:5] + handled(b.stderr)) == len(hf)
