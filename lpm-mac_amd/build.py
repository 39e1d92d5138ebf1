"""Build liblpm_mi355x.so in-tree for gfx950 (hipcc for the kernels, g++ for host setup code)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblpm_mi355x.so")
ARCH = os.environ.get("LPM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
# per-file flags: no FMA contraction where FP64 results must follow the reference's operation
# order; host OpenMP for the dump writer's parallel formatting
EXTRA = {"kernels_phys.hip": ["-ffp-contract=off"], "ctx.hip": ["-ffp-contract=off"], "io.hip": ["-fopenmp"]}


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in ("kernels_phys.hip", "kernels_cg.hip", "kernels_coarse.hip", "kernels_adj.hip", "ctx.hip", "io.hip"):
        o = os.path.join(OBJ, src + ".o")
        run(["hipcc", "-c", *HIP_FLAGS, *EXTRA.get(src, []), "-o", o, os.path.join(CSRC, src)])
        objs.append(o)
    o = os.path.join(OBJ, "topology.o")
    run(["g++", "-c", "-O2", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
         "-o", o, os.path.join(CSRC, "topology.cpp")])
    objs.append(o)
    tmp = LIB + ".tmp"
    run(["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
