"""Compile the gfx950 HIP sources into naivedynamics.jl_amd/libnaivehip.so.

Every translation unit that feeds a bit-exact comparison with the reference
(keys, sort, tree, traversal) is compiled with -ffp-contract=off so no
multiply-add is fused; the force/integration units allow contraction.  All
units keep IEEE single-precision division/sqrt and FP32 denormals.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnaivehip.so")
ARCH = os.environ.get("ND_OFFLOAD_ARCH", "gfx950")

EXACT = ["nd_ctx.hip", "nd_morton.hip", "nd_sort.hip", "nd_lbvh.hip", "nd_traverse.hip",
         "nd_cells.hip"]
FAST = ["nd_force.hip", "nd_sim.hip", "nd_adjoint.hip"]

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def _compile(src, contract):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [hipcc()] + COMMON + ["-ffp-contract=" + contract, "-c", os.path.join(CSRC, src),
                                "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = [(s, "off") for s in EXACT] + [(s, "fast") for s in FAST]
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
