"""Build libstreet_solve.so (HIP, gfx950) in-tree: poker-solver_amd/build/street_solve.so.

    python poker-solver_amd/build_lib.py [--jobs N]

Every source is compiled by hipcc for gfx950 with packed-FP32 instruction
selection disabled, then linked into one shared library exporting the C-ABI
of include/street_solve.h, include/street_solve_leaf.h and
include/street_solve_range.h (and the test entry points of
include/street_solve_debug.h).
"""

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "build")
OBJ = os.path.join(BUILD, "obj")
LIB = os.path.join(BUILD, "street_solve.so")

SOURCES = ["ss_tree.cpp", "ss_runtime.hip", "ss_kernels.hip", "ss_precision.hip", "ss_capi.hip", "ss_server.hip", "ss_leaf.hip", "ss_range.hip", "ss_flop.hip"]

ARCH = os.environ.get("SS_OFFLOAD_ARCH", "gfx950")


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found (ROCm is required to build the library)")
    return p


def compile_flags():
    return ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", os.path.join(REPO, "include"), "-I", CSRC]


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()

    def one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + compile_flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(SOURCES))) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
