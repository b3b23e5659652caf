"""Build libregot_hip.so (HIP kernels + host drivers) for gfx950 with hipcc.

Usage: python regot-python_amd/build.py [--debug]
The library lands in regot-python_amd/regot/ next to the Python package that
loads it.  Every HIP translation unit is compiled with packed FP32 math
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "regot", "libregot_hip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["dense_kernels.hip", "vec_kernels.hip", "sparse_kernels.hip", "hessian.hip",
           "problem.hip", "batch.hip", "cost_kernels.hip", "grad.hip", "solvers.cpp",
           "newton.cpp", "first_order.cpp", "apdagd_kernels.hip", "objgrad_kernels.hip",
           "primal_kernels.hip", "bcd_fused.hip", "qrot_kernels.hip", "qrot.cpp", "noalloc.hip",
           "hess_mf.hip"]
ARCH = os.environ.get("REGOT_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def flags(debug=False):
    f = ["-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
         "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
    f += ["-O0", "-g"] if debug else ["-O3"]
    return f


def build(debug=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    fl = flags(debug)

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + fl + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s%s" % (src, r.stdout, r.stderr))
        return obj

    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--debug", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.debug, a.verbose))
