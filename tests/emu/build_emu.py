"""Builds tests/_build/emu/libhydist_emu.so: the library's sources compiled
for the CPU against tests/emu/hip/hip_runtime.h (test infrastructure).

Kernel launches `k<<<grid, block, lds, stream>>>(args)` are rewritten to
`hd_emu::launcher("k", k, grid, block, lds, stream)(args)`, and the one
`extern __shared__` array to the launch's dynamic LDS buffer.  Built with
-ffp-contract=off like the HIP library and the oracle, so results can be
compared bit for bit."""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from hydist_c_amd import build as hb  # noqa: E402

OUT = os.path.join(ROOT, "tests", "_build", "emu")
LIB = os.path.join(OUT, "libhydist_emu.so")
LAUNCH = re.compile(r"([A-Za-z_][\w:]*(?:<[^<>;(){}]*>)?)\s*<<<(.*?)>>>\s*\(", re.S)


def transform(text):
    text = re.sub(r"extern __shared__ (double|unsigned char) (\w+)\[\];", r"\1 *\2 = (\1 *)hd_emu::dyn_lds();", text)
    return LAUNCH.sub(lambda m: f'hd_emu::launcher("{m.group(1)}", {m.group(1)}, {m.group(2)})(', text)


def _compile(args):
    src, obj = args
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-w",
           "-I" + os.path.join(ROOT, "tests", "emu"), "-I" + os.path.join(ROOT, "include"), "-I" + hb.CSRC,
           "-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emu build failed for {src}:\n{r.stderr[-4000:]}")
    return obj


def build(verbose=False):
    src_dir = os.path.join(OUT, "src")
    os.makedirs(src_dir, exist_ok=True)
    jobs = []
    for s in hb.SOURCES:
        p = os.path.join(hb.CSRC, s)
        if not os.path.exists(p):
            continue
        dst = os.path.join(src_dir, os.path.splitext(s)[0] + ".cpp")
        with open(p) as f:
            text = transform(f.read())
        with open(dst, "w") as f:
            f.write(text)
        jobs.append((dst, os.path.join(OUT, os.path.splitext(s)[0] + ".o")))
    jobs.append((os.path.join(HERE, "emu_runtime.cpp"), os.path.join(OUT, "emu_runtime.o")))
    with ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_compile, jobs))
    r = subprocess.run(["g++", "-shared", "-o", LIB] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-4000:])
    drv = os.path.join(OUT, "compat_driver")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-I" + os.path.join(ROOT, "tests", "emu"),
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "compat", "compat_driver.cpp"), "-o", drv,
                        "-L" + OUT, "-lhydist_emu", "-Wl,-rpath," + OUT], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-4000:])
    if verbose:
        print("built", LIB, drv)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
