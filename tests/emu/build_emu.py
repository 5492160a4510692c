"""Builds tests/_build/emu/libhydist_emu.so: the library's sources compiled
for the CPU against tests/emu/hip/hip_runtime.h (test infrastructure).

Kernel launches `k<<<grid, block, lds, stream>>>(args)` are rewritten to
`hd_emu::launcher("k", k, grid, block, lds, stream)(args)`, and the one
`extern __shared__` array to the launch's dynamic LDS buffer.  Built with
-ffp-contract=off like the HIP library and the oracle, so results can be
compared bit for bit."""
import fcntl
import glob
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
# the same with AddressSanitizer on the host code (the library's own C++:
# contexts, the compat layer's loaders, the serving loop's threads); the
# emulated device buffers are mmap'd and stay checked by their guard pages
OUT_ASAN = os.path.join(ROOT, "tests", "_build", "emu_asan")
ASAN = ["-fsanitize=address", "-fno-omit-frame-pointer"]
# and with ThreadSanitizer (the serving loop's worker, its helper thread
# and the API calls of client threads; the emulator announces its fibers)
OUT_TSAN = os.path.join(ROOT, "tests", "_build", "emu_tsan")
TSAN = ["-fsanitize=thread", "-fno-omit-frame-pointer"]
# and with UndefinedBehaviorSanitizer, which here also covers the kernels
# (compiled as C++ for the CPU): integer overflow in index arithmetic,
# out-of-range shifts, misaligned or null accesses, invalid bool and enum
# loads; the first finding aborts
OUT_UBSAN = os.path.join(ROOT, "tests", "_build", "emu_ubsan")
UBSAN = ["-fsanitize=undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
VARIANTS = {False: (OUT, []), True: (OUT_ASAN, ASAN), "address": (OUT_ASAN, ASAN), "thread": (OUT_TSAN, TSAN),
            "undefined": (OUT_UBSAN, UBSAN)}
LAUNCH = re.compile(r"([A-Za-z_][\w:]*(?:<[^<>;(){}]*>)?)\s*<<<(.*?)>>>\s*\(", re.S)


def transform(text):
    text = re.sub(r"extern __shared__ (double|unsigned char) (\w+)\[\];", r"\1 *\2 = (\1 *)hd_emu::dyn_lds();", text)
    return LAUNCH.sub(lambda m: f'hd_emu::launcher("{m.group(1)}", {m.group(1)}, {m.group(2)})(', text)


def _compile(args, extra=()):
    src, obj = args
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-w", *extra,
           "-I" + os.path.join(ROOT, "tests", "emu"), "-I" + os.path.join(ROOT, "include"), "-I" + hb.CSRC,
           "-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emu build failed for {src}:\n{r.stderr[-4000:]}")
    return obj


def _inputs():
    """Every file a build reads: the library's sources and headers, the
    emulated runtime, the compat driver and this script."""
    files = glob.glob(os.path.join(hb.CSRC, "*")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    files += glob.glob(os.path.join(HERE, "*.cpp")) + glob.glob(os.path.join(HERE, "hip", "*.h"))
    files += [os.path.join(ROOT, "tests", "compat", "compat_driver.cpp"), os.path.abspath(__file__)]
    return [f for f in files if os.path.isfile(f)]


def build(verbose=False, sanitize=False, force=False):
    """Builds unless the outputs are newer than every input (force: always);
    concurrent callers (test workers) wait on a lock and build once."""
    if sanitize not in VARIANTS:
        raise ValueError(f"sanitize: {sanitize!r}")
    out = VARIANTS[sanitize][0]
    os.makedirs(out, exist_ok=True)
    lib = os.path.join(out, "libhydist_emu.so")
    drv = os.path.join(out, "compat_driver")
    with open(os.path.join(out, ".lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        newest = max(os.path.getmtime(f) for f in _inputs())
        if not force and all(os.path.exists(f) and os.path.getmtime(f) > newest for f in (lib, drv)):
            if verbose:
                print("up to date", lib, drv)
            return lib
        return _build(verbose, sanitize)


def _build(verbose=False, sanitize=False):
    """Builds the emulated library and the compat driver against it into
    tests/_build/emu (sanitize True or "address": with AddressSanitizer,
    into tests/_build/emu_asan; "thread": with ThreadSanitizer, into
    tests/_build/emu_tsan; "undefined": with UndefinedBehaviorSanitizer,
    into tests/_build/emu_ubsan).  Returns the library's path."""
    out, extra = VARIANTS[sanitize]
    lib = os.path.join(out, "libhydist_emu.so")
    src_dir = os.path.join(out, "src")
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
        jobs.append((dst, os.path.join(out, os.path.splitext(s)[0] + ".o")))
    jobs.append((os.path.join(HERE, "emu_runtime.cpp"), os.path.join(out, "emu_runtime.o")))
    with ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(lambda j: _compile(j, extra), jobs))
    r = subprocess.run(["g++", "-shared", *extra, "-o", lib] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-4000:])
    drv = os.path.join(out, "compat_driver")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", *extra, "-I" + os.path.join(ROOT, "tests", "emu"),
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "compat", "compat_driver.cpp"), "-o", drv,
                        "-L" + out, "-lhydist_emu", "-Wl,-rpath," + out], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-4000:])
    if verbose:
        print("built", lib, drv)
    return lib


if __name__ == "__main__":
    build(verbose=True, force="--force" in sys.argv)
