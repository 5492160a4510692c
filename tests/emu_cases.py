"""Failure-handling cases run against the CPU-emulated build of the library
(tests/emu): `python tests/emu_cases.py <case>` prints one JSON line with
what the C ABI returned.  Test infrastructure; tests/test_failure_emu.py
runs each case in its own process so the emulated library never meets the
GPU build in one interpreter."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import emu_lib  # noqa: E402

emu_lib.use_emulator()

import numpy as np  # noqa: E402

from hydist_c_amd import engine, scenarios as S  # noqa: E402


def rc_of(fn, *a):
    try:
        fn(*a)
        return 0
    except engine.HydistError as e:
        return int(str(e).split()[2].rstrip(":"))


def wet_cell(sim):
    kh = sim.get("KHOUOT")
    ii, jj = np.nonzero(kh[2:-2, 2:-2] == 0)
    return ii[len(ii) // 2] + 2, jj[len(jj) // 2] + 2


def case_watchdog():
    sim = engine.Simulation(S.open_box(24, 18, seed=3))
    sim.init()
    out = {}
    sim.set_watchdog(1)
    out["clean_run"] = rc_of(sim.run, 6.0)
    out["clean_sync"] = rc_of(sim.sync)
    z = sim.get("Z")
    i, j = wet_cell(sim)
    z[i, j] = np.nan
    sim.set("Z", z)
    sim.set_watchdog(2)
    t0 = sim.t
    out["poisoned_run"] = rc_of(sim.run, 40.0)
    out["steps_before_stop"] = int(round((sim.t - t0) / 2.0))
    out["sync_after"] = rc_of(sim.sync)
    out["run_after"] = rc_of(sim.run, 60.0)
    sim.set_watchdog(0)
    out["sync_off"] = rc_of(sim.sync)
    return out


def case_watchdog_step():
    sim = engine.Simulation(S.open_box(24, 18, seed=3))
    sim.init()
    sim.set_watchdog(3)
    v = sim.get("V")
    i, j = wet_cell(sim)
    v[i, j] = np.inf
    sim.set("V", v)
    out = {"step1": rc_of(sim.step, 1), "step2": rc_of(sim.step, 1)}
    out["sync_before_check"] = rc_of(sim.sync)
    out["step3"] = rc_of(sim.step, 1)
    out["sync_after_check"] = rc_of(sim.sync)
    return out


def case_fault():
    sim = engine.Simulation(S.open_box(24, 18, seed=3))
    sim.init()
    out = {}
    sim.inject_fault(1)
    out["uz"] = rc_of(sim.halfstep_uz, 1.0)
    out["step_uz"] = rc_of(sim.step, 1)
    sim.inject_fault(2)
    out["uz_ok"] = rc_of(sim.halfstep_uz, 1.0)
    out["vz"] = rc_of(sim.halfstep_vz, 2.0)
    sim.inject_fault(0)
    out["step_ok"] = rc_of(sim.step, 1)
    out["bad_stage"] = rc_of(sim.inject_fault, 3)
    return out


if __name__ == "__main__":
    print(json.dumps(globals()["case_" + sys.argv[1]]()), flush=True)
