"""Build libpda.so (HIP, gfx950) in-tree with hipcc.

Usage: python pressio-demoapps_amd/build.py [--jobs N]
The library lands next to this file as libpda.so; nothing is installed.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpda.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PDA_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no packed FP32 instructions in any code object
COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

SOURCES = ["pda_mesh.cpp", "pda_problem.cpp", "pda_kernels.hip", "pda_box_kernels.hip",
           "pda_stepper.hip", "pda_implicit.hip",
           "pda_guessnet.hip", "pda_batch.hip", "pda_noalloc.hip"]


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    srcp = os.path.join(CSRC, src)
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", srcp, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass warns that it does not know -packed-fp32-ops; drop that line
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "unknown target feature" not in l.lower())
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return out


def build(jobs=None):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    print(build(a.jobs))
