"""One simulated hour of the 2048 x 1024 meander + islands workload
(or, with `c3`, 600 s of the 4096 x 4096 braided river at dt = 1 s)
(sediment and FP64 bed change from t = 300 s, dt = 2 s, 1800 steps) in each
GPU form, compared with the CPU oracle after the hour: max |dZ|, |dU|,
|dV|, |dFS|, |d dH| over cells wet in both, and the wet-mask difference.
The fast forms' targets: |dZ| <= 1 mm, |dU|, |dV| <= 1e-4 m/s.

Run directly (not collected by pytest):  python tests/accuracy_one_hour.py
[steps [M N]] (a smaller meander + islands mesh for a quick check)
Prints progress lines and one JSON line with the table."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import oracle_lib as ol  # noqa: E402
from hydist_c_amd import abi, engine, scenarios as S  # noqa: E402
from parity_util import bits_equal  # noqa: E402

FORMS = {"plain": abi.HD_FORM_FP64_PLAIN, "fast": abi.HD_FORM_FP64_FAST, "fp32": abi.HD_FORM_FP32_REFINE}


def log(*a):
    print(*a, flush=True)


def main():
    # arguments: [steps] [M N] | c3 [steps]
    if len(sys.argv) > 1 and sys.argv[1] == "c3":
        # 4096 x 4096 braided river, dt = 1 s, 600 s, hydraulics only
        sc = S.config3()
        steps = int(sys.argv[2]) if len(sys.argv) > 2 else 600
        sed = False
    else:
        steps = int(sys.argv[1]) if len(sys.argv) > 1 else 1800
        sc = S.config2() if len(sys.argv) < 3 else S.meander_islands(int(sys.argv[2]), int(sys.argv[3]), seed=2)
        sed = True
    log(f"[acc] mesh {sc.M}x{sc.N}, {steps} steps")
    sims = {}
    for name, form in FORMS.items():
        t0 = time.time()
        g = engine.Simulation(sc, form=form, sediment=sed, bed_change=sed)
        g.init()
        for k in range(0, steps, 300):
            g.step(min(300, steps - k))
            g.sync()
        sims[name] = g
        log(f"[acc] gpu {name}: {time.time() - t0:.1f} s, t = {g.t}")
    # the oracle seeded with the GPU's pow/log-derived init fields
    o = ol.Oracle(sc, cfg=ol.make_config(sediment=sed, bed_change=sed))
    o.init()
    ref = engine.Simulation(sc, form=abi.HD_FORM_FP64_PLAIN, sediment=sed, bed_change=sed)
    ref.init()
    for f in ("KX1", "KY1"):
        o.set(f, ref.get(f))
    o.set_boundary("HI", ref.get_boundary("HI"))
    o.set_hmax(ref.get_hmax())
    ref.close()
    ol.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1)
    t0 = time.time()
    for k in range(0, steps, 100):
        o.step(min(100, steps - k))
        log(f"[acc] oracle step {min(k + 100, steps)}/{steps} ({time.time() - t0:.0f} s)")
    table = {}
    kh_o = o.get("KHOUOT")
    for name, g in sims.items():
        kh = g.get("KHOUOT")
        wet = (kh == 0) & (kh_o == 0)
        row = {"mask_diff_cells": int((kh != kh_o).sum())}
        for f in ("Z", "U", "V", "FS", "DH"):
            a, b = g.get(f), o.get(f)
            row[f"max_abs_d{f}"] = float(np.max(np.abs(a[wet] - b[wet]))) if wet.any() else 0.0
            row[f"bitwise_{f}"] = bool(bits_equal(a, b))
        table[name] = row
        log(f"[acc] {name}: {row}")
    print(json.dumps({"steps": steps, "simulated_s": steps * sims["plain"].coeffs.dT, "mesh": [sc.M, sc.N],
                      "forms": table}), flush=True)


if __name__ == "__main__":
    main()
