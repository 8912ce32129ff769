"""Build libkmx.so (gfx950) in-tree: dg-sparsegrid_amd/lib/libkmx.so.

    python dg-sparsegrid_amd/build.py [-j N] [--verbose]

Each .hip source is compiled to an object in parallel with hipcc, then linked
into one shared library.  Packed FP32 instructions are disabled for all device
code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libkmx.so")
SOURCES = ["kmx_api.hip", "kmx_apply.hip", "kmx_apply_f64.hip", "kmx_apply_f32.hip", "kmx_kron.hip", "kmx_blocks.hip", "kmx_realspace.hip", "kmx_sweep.hip", "kmx_coeff.hip", "kmx_gmres.hip", "kmx_assemble.hip"]
ARCH = "gfx950"


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else (shutil.which("hipcc") or "hipcc")


COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def compile_one(src, verbose):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(jobs=5, verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the plain-C example caller of the ABI (host code only)
    demo = os.path.join(LIBDIR, "kmx_demo")
    subprocess.run(["gcc", "-O2", "-o", demo, os.path.join(HERE, "examples", "kmx_demo.c"), "-L" + LIBDIR, "-lkmx",
                    "-Wl,-rpath,$ORIGIN", "-lm"], check=True)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=5)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.verbose))


if __name__ == "__main__":
    sys.exit(main())
