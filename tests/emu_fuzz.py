"""Parity fuzzing on the CPU (test infrastructure, not collected by
pytest): `python tests/emu_fuzz.py FIRST LAST` draws the scenarios
FIRST..LAST-1 of the randomised parity distribution
(test_gpu_random_parity.random_scenario), runs each in every form through
the CPU-emulated library (tests/emu) with line solvers chosen at random
per sweep, and against the oracle, and prints one JSON line per run
("BAD" where a bar is missed).  Bars: plain bit for bit on the hydraulic
fields and the mask; the fast forms at the randomised test's bars, scaled
by the field's largest magnitude when it exceeds 1 (a draw whose flow
runs away grows the values, and rounding grows with them)."""
import json
import os
import sys
import traceback
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def work(k):
    os.environ["OMP_NUM_THREADS"] = "1"
    sys.path.insert(0, ROOT)
    sys.path.insert(0, HERE)
    import emu_lib
    emu_lib.use_emulator()
    import numpy as np
    from hydist_c_amd import engine
    import test_gpu_random_parity as R
    from parity_util import HYDRO, bits_equal, make_pair

    rng = np.random.default_rng(k)
    sc, sed = R.random_scenario(k)
    out = []
    for form in (R.PLAIN, R.FAST, R.FP32):
        try:
            g, o = make_pair(sc, form=form, sediment=sed, bed_change=sed)
            chosen = []
            if form == R.PLAIN:
                kinds = [(engine.SOLVER_DEFAULT, 0), (engine.SOLVER_SEGMENTS, 0)]
            else:
                kinds = [(engine.SOLVER_PARTITION, 2), (engine.SOLVER_PARTITION, 8), (engine.SOLVER_CR, 64),
                         (engine.SOLVER_CR, 512)] + ([(engine.SOLVER_SEGMENTS, 0)] if form == R.FAST else [])
            for d in (0, 1):
                for hyd in (0, 1):
                    kind, p = kinds[int(rng.integers(len(kinds)))]
                    try:
                        engine.set_solver(g, d, hyd, kind, p)
                        chosen.append((d, hyd, kind, p))
                    except engine.HydistError:  # a line too long for this choice keeps the default
                        pass
            steps = int(rng.integers(3, 12))
            g.step(steps)
            o.step(steps)
            missed = {}
            if not np.array_equal(g.get("KHOUOT"), o.get("KHOUOT")):
                missed["KHOUOT"] = "differs"
            for f in HYDRO:
                a, b = g.get(f), o.get(f)
                if form == R.PLAIN:
                    if not bits_equal(a, b):
                        missed[f] = "bits"
                    continue
                if not np.array_equal(np.isnan(a), np.isnan(b)):
                    missed[f] = "nan pattern"
                    continue
                fin = np.isfinite(a) & np.isfinite(b)
                if fin.any():
                    err = float(np.max(np.abs(a[fin] - b[fin])))
                    if err > R.HYD_TOL[form] * max(1.0, float(np.max(np.abs(b[fin])))):
                        missed[f] = err
            if sed:
                for f in ("FS", "DH"):
                    e = R.rel(g.get(f), o.get(f))
                    if e > R.REL_TOL[form]:
                        missed[f] = e
            g.close()
            out.append({"draw": k, "scenario": sc.name, "form": R.NAMES[form], "steps": steps,
                        "solvers": chosen, "missed": missed})
        except Exception:
            out.append({"draw": k, "scenario": sc.name, "form": R.NAMES[form], "error": traceback.format_exc()[-800:]})
    return out


if __name__ == "__main__":
    lo, hi = int(sys.argv[1]), int(sys.argv[2])
    bad = 0
    with ProcessPoolExecutor(4) as ex:
        for res in ex.map(work, range(lo, hi)):
            for r in res:
                miss = "error" in r or r["missed"]
                bad += bool(miss)
                print(("BAD " if miss else "ok  ") + json.dumps(r), flush=True)
    print(f"{bad} runs missed a bar", flush=True)
    sys.exit(1 if bad else 0)
