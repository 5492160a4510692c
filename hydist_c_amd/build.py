"""Builds the gfx950 library (hydist_c_amd/libhydist_amd.so) with hipcc.

Every HIP translation unit is compiled with -ffp-contract=off (so FP64
results built from + - * / sqrt match the CPU oracle bit for bit) and with
the packed-FP32 target feature disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhydist_amd.so")

SOURCES = ["hd_kernels.hip", "hd_sweep.hip", "hd_api.hip", "hd_context.hip", "hd_server.hip",
           "hd_compat.hip", "hd_gtsv.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_compat_driver(verbose=False):
    """tests/compat/compat_driver: a program using the reference's host
    surface (include/hydist_compat.h) linked against the library."""
    src = os.path.join(ROOT, "tests", "compat", "compat_driver.cpp")
    out_dir = os.path.join(ROOT, "tests", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "compat_driver")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", exe,
           "-L" + PKG, "-lhydist_amd", "-Wl,-rpath,$ORIGIN/../../hydist_c_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compat driver build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", exe)
    return exe

def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_compat_driver(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
