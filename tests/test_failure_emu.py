"""Failure handling through the C ABI, on the CPU: the NaN/Inf watchdog
(hd_set_watchdog) and fault injection (hd_inject_fault) against the
emulated build of the library (tests/emu), each case in its own process
(tests/emu_cases.py)."""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
E_NUMERIC, E_INVALID, E_FAULT = -4, -1, -7


@pytest.fixture(scope="module")
def emu_built():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    build_emu.build()


def run_case(name):
    r = subprocess.run([sys.executable, os.path.join(HERE, "emu_cases.py"), name], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_watchdog_stops_run_at_first_check(emu_built):
    out = run_case("watchdog")
    assert out["clean_run"] == 0 and out["clean_sync"] == 0
    # poisoned z: the check after the 2nd step finds it and stops the run
    assert out["poisoned_run"] == E_NUMERIC
    assert out["steps_before_stop"] == 2
    # the finding sticks until the watchdog is set again
    assert out["sync_after"] == E_NUMERIC and out["run_after"] == E_NUMERIC
    assert out["sync_off"] == 0


def test_watchdog_after_step_reported_by_sync(emu_built):
    out = run_case("watchdog_step")
    assert out["step1"] == 0 and out["step2"] == 0
    assert out["sync_before_check"] == 0   # no check has run yet (every 3 steps)
    assert out["step3"] == 0               # hd_step queues the check, does not wait
    assert out["sync_after_check"] == E_NUMERIC


def test_injected_faults(emu_built):
    out = run_case("fault")
    assert out["uz"] == E_FAULT and out["step_uz"] == E_FAULT
    assert out["uz_ok"] == 0 and out["vz"] == E_FAULT
    assert out["step_ok"] == 0
    assert out["bad_stage"] == E_INVALID
