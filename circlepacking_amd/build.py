"""Build step of the lane-packing prover: compiles libcpl_hip.so for gfx950.

    python -m circlepacking_amd.build          # HIP library only
The library is built in-tree (circlepacking_amd/libcpl_hip.so) so that it
travels with the source tree to the machine that runs it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcpl_hip.so")
SOURCES = ["cpl_hip.hip", "cpl_mlp.hip"]
HEADERS = ["cpl_round.hpp", "cpl_iv.hpp", "cpl_asin.hpp", "cpl_lane.hpp", "cpl_kernels.hpp", "cpl_mlp.hpp",
           "cpl_host.hpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no implicit contraction: every fused multiply-add is written explicitly
    "-ffp-contract=off",
    "-fno-fast-math",
    # packed FP32 instructions are not used on this hardware pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-function",
    # the scheduler's instruction-level-parallelism strategy: the interval
    # criteria are long dependency chains at three to four waves per SIMD
    # (graph and plain forms 2-2.5 %, reference-grid kernels 1-2 % faster
    # than the default occupancy strategy, identical results)
    "-mllvm", "-amdgpu-sched-strategy=max-ilp",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build_hip(verbose=False):
    """Compiles every source to an object in parallel, then links the library."""
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_obj", src.replace(".hip", ".o"))
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + [
            "-I", os.path.join(ROOT, "include"), "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    noise = "is not a recognized feature for this target"
    for src, p in procs:
        _, stderr = p.communicate()
        err = "\n".join(l for l in stderr.splitlines() if noise not in l)
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src} ({p.returncode}):\n{err}")
        if err.strip() and verbose:
            print(err, file=sys.stderr)
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed ({res.returncode}):\n{res.stderr}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


PROVER = os.path.join(HERE, "cpl_prover")


def build_driver(verbose=False):
    """The C++ prover executable (prover.cpp), linked against libcpl_hip.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", PROVER, os.path.join(CSRC, "prover.cpp"),
           "-I", os.path.join(ROOT, "include"), "-L", HERE, "-lcpl_hip", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed ({res.returncode}):\n{res.stderr}")
    return PROVER


COST_PROBE = os.path.join(HERE, "cpl_cost_probe")


def build_cost_probe(verbose=False, occupancy=None, out=None):
    """The issue-cost probe of the criteria (cost_probe.hip), a standalone HIP
    program; occupancy: waves per SIMD of its criteria throughput kernel."""
    out = out or COST_PROBE
    cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")] + [
        "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-o", out + ".tmp", os.path.join(CSRC, "cost_probe.hip")]
    if occupancy:
        cmd.insert(1, f"-DCPL_PROBE_OCC={int(occupancy)}")
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on cost_probe.hip ({res.returncode}):\n{res.stderr}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build_hip(verbose=True))
    print(build_driver(verbose=True))
    print(build_cost_probe(verbose=True))
