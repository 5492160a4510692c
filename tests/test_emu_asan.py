"""The library under AddressSanitizer, on the CPU: the emulated build
(tests/emu) compiled with -fsanitize=address runs the reference-surface
driver (compat layer: input loading, contexts, the time loop,
save_result), the serving loop with a batched group, a context made on
the helper thread and a cancelled run, and a cross-section of the GPU
cases, whose kernels are C++ for the CPU here, so a per-lane array indexed
out of its bounds is caught too.  A heap or stack error ends the child
with ASan's report.  (The emulated device buffers stay checked by their
guard pages; the drop-in path's context lives for the process, as the
reference's allocations do, so leak checks are off.)"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SERVER = r"""
import sys
sys.path.insert(0, {root!r})
from hydist_c_amd import abi, engine, scenarios as S
engine.LIB_PATH = {lib!r}
engine._lib = None
FAST = abi.HD_FORM_FP64_FAST
big = S.open_box(64, 40, seed=5)
small = [S.open_box(24, 18, seed=20 + k) for k in range(3)]
srv = engine.Server(hbm_budget=engine.footprint(big) + 1, max_active=8, steps_per_quantum=2, use_graph=False)
ids = [srv.submit(big, 6.0, form=FAST)] + [srv.submit(sc, 8.0, form=FAST, sediment=True, bed_change=True)
                                            for sc in small]
gone = srv.submit(S.open_box(24, 18, seed=99), 200.0)
srv.cancel(gone)
assert [srv.wait(r, timeout=300.0) for r in ids] == [engine.RUN_DONE] * 4
st = srv.stats()
assert st["batched"] == 3 and st["cancelled"] == 1, st
for r in ids[1:]:
    srv.result(r, "Z")
srv.close()
print("server ok", flush=True)
"""


@pytest.fixture(scope="module")
def asan_build():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    lib = build_emu.build(sanitize=True)
    rt = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(rt) and os.path.exists(rt), rt
    return lib, os.path.join(os.path.dirname(lib), "compat_driver"), rt


def test_compat_driver_under_asan(asan_build, tmp_path):
    from hydist_c_amd import inputs, scenarios as S
    _, drv, _ = asan_build
    d = inputs.write_scenario_inputs(S.open_box(24, 18, types=(2, 0, 0, 0)), str(tmp_path / "Inputs"))
    for f in ("U", "V", "Z", "FS"):
        os.makedirs(tmp_path / "Outputs" / f, exist_ok=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([drv, "-f", d, "-h", "0", "-m", "0", "-s", "10", "-i", "4"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert os.path.exists(tmp_path / "Outputs" / "Z" / "z_8.txt")


def test_serving_loop_under_asan(asan_build):
    lib, _, rt = asan_build
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, "-c", SERVER.format(root=ROOT, lib=lib)], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0 and "server ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]


# kernels of every solver and form, the tridiagonal entry points and a
# degenerate mesh
CASES = " or ".join([
    "gtsv", "noalloc", "(random_scenario and [0-)", "(thomas_per_system and [box])",
    "(test_plain_fp64_bitwise and 12_0)", "test_fast_form_sediment",
    "(test_every_line_solver_close_to_oracle and part4)", "(edge_mesh_fast and onecell)",
])


def test_gpu_cases_under_asan(asan_build):
    import test_emu_suite as suite
    lib, _, rt = asan_build
    suite.run_cases(CASES, 15, HD_EMU_LIB=lib, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
