"""The library under UndefinedBehaviorSanitizer, on the CPU: the emulated
build (tests/emu) compiled with -fsanitize=undefined, where the kernels
are C++ for the CPU too, so their index arithmetic (signed overflow,
shifts), accesses (null, misaligned) and loads (bool, enum) are checked
along with the host code.  A cross-section of the GPU cases runs against
it (the first finding aborts the child), and the reference-surface driver
runs a short real-input case."""
import os
import subprocess
import sys

import pytest

import test_emu_suite as suite

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ubsan_build():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    lib = build_emu.build(sanitize="undefined")
    return lib, os.path.join(os.path.dirname(lib), "compat_driver")


def test_gpu_cases_under_ubsan(ubsan_build):
    lib, _ = ubsan_build
    suite.run_cases(suite.REVERSE, 20, HD_EMU_LIB=lib, UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def test_compat_driver_under_ubsan(ubsan_build, tmp_path):
    from hydist_c_amd import inputs, scenarios as S
    _, drv = ubsan_build
    d = inputs.write_scenario_inputs(S.open_box(24, 18, types=(2, 0, 0, 0)), str(tmp_path / "Inputs"))
    for f in ("U", "V", "Z", "FS"):
        os.makedirs(tmp_path / "Outputs" / f, exist_ok=True)
    env = dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", HYDIST_FORM="fast")
    r = subprocess.run([drv, "-f", d, "-h", "0", "-m", "0", "-s", "10", "-i", "4"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
    assert os.path.exists(tmp_path / "Outputs" / "Z" / "z_8.txt")
