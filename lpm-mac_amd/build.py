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


def build_autograd(verbose=False):
    """The torch autograd harness (host/lpm_autograd.cpp) as the Python module lpm_autograd next to
    the library, linked to it through its C ABI. Host code only (g++ against torch's headers)."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce

    tlib = os.path.join(os.path.dirname(torch.__file__), "lib")
    out = os.path.join(HERE, "lpm_autograd" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out + ".tmp"
    cmd = ["g++", "-shared", "-fPIC", "-O2", "-std=c++17", f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=lpm_autograd",
           *[f"-I{p}" for p in ce.include_paths()], "-I", sysconfig.get_paths()["include"], "-I", os.path.join(ROOT, "include"),
           "-o", tmp, os.path.join(HERE, "host", "lpm_autograd.cpp"), f"-L{HERE}", "-llpm_mi355x", f"-L{tlib}", "-lc10",
           "-ltorch", "-ltorch_cpu", "-ltorch_python", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tlib}"]
    run(cmd)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_autograd(verbose=True)
