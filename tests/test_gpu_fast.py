"""GPU fast forms against the CPU oracle.

HD_FORM_FP64_FAST solves the z-only (Schur) system of each line in FP64 by
the partition solver (or, when chosen, cyclic reduction);
HD_FORM_FP32_REFINE solves it in FP32 and adds one FP64
residual-refinement solve.  Both differ from the reference's interleaved
Thomas only by round-off, so they are compared with tolerances:

  FP64_FAST:   |dZ|, |dU|, |dV| <= 1e-13; FS, dH within 1e-12 relative
  FP32_REFINE: |dZ|, |dU|, |dV| <= 5e-10; FS, dH within 1e-9 relative
  wet/dry mask and segment tables: identical
(the tidal-flats case keeps its own bar, explained there; the one-hour
run of config 1 is held to the accuracy bar of the mixed form).
"""
import numpy as np
import pytest

from hydist_c_amd import engine, scenarios as S
from hydist_c_amd.abi import HD_FORM_FP32_REFINE, HD_FORM_FP64_FAST
from parity_util import make_pair

pytestmark = pytest.mark.gpu

# hydraulics (m, m/s) and FS / bed change relative to the field's largest
# value: about 100 times the largest error observed on these cases (fast
# FP64 ~1e-15, FP32 + refine a few 1e-12 on the hydraulics)
TOL = {HD_FORM_FP64_FAST: 1e-13, HD_FORM_FP32_REFINE: 5e-10}
REL_TOL = {HD_FORM_FP64_FAST: 1e-12, HD_FORM_FP32_REFINE: 1e-9}
NAMES = {HD_FORM_FP64_FAST: "fp64_fast", HD_FORM_FP32_REFINE: "fp32_refine"}


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
@pytest.mark.parametrize("maker,steps", [
    (lambda: S.open_box(40, 28), 10),
    (lambda: S.open_box(57, 33, seed=9, types=(0, 2, 0, 2)), 10),
    (lambda: S.config1(), 30),
    (lambda: S.meander_islands(160, 96, n_islands=8), 20),
])
def test_fast_form_close_to_oracle(form, maker, steps):
    sc = maker()
    print(f"[case] fast {NAMES[form]} {sc.name} steps={steps}", flush=True)
    g, o = make_pair(sc, form=form)
    g.step(steps)
    o.step(steps)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("Z", "U", "V", "HMOI"):
        d = np.max(np.abs(g.get(f) - o.get(f)))
        print(f"[{f}] max abs error {d:.3e}", flush=True)
        assert d <= TOL[form], (f, d)


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
def test_fast_form_wetting_drying(form):
    """Tidal flats: cells dry (33 of them change state in the first 100
    steps) and the upwind switches on H_moi <= H_TINH
    (UVZSolver_multithread.cpp:186-201, 252-268) act at the front.
    Bar after 100 steps: identical masks, |dZ|, |dU|, |dV| <= 1e-6.
    The run stops at 100 steps on purpose: from about step 110 the front
    of this case produces velocity spikes in the FP64 oracle itself (|U|
    of several m/s on films of a few cm), where differences of round-off
    size grow tenfold every few steps for every solver, the oracle's own
    Thomas included against any reordering of it."""
    sc = S.tidal_flats(64, 40)
    print(f"[case] fast-flats {NAMES[form]}", flush=True)
    g, o = make_pair(sc, form=form)
    k0 = o.get("KHOUOT").copy()
    g.step(100)
    o.step(100)
    kg, ko = g.get("KHOUOT"), o.get("KHOUOT")
    assert int((ko != k0).sum()) >= 20, "the case no longer dries cells"
    assert np.array_equal(kg, ko)
    for f in ("Z", "U", "V"):
        d = np.max(np.abs(g.get(f) - o.get(f)))
        print(f"[flats] {NAMES[form]} max|d{f}| = {d:.3e}", flush=True)
        assert d <= 1e-6, (f, d)


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
def test_fast_form_sediment(form):
    sc = S.open_box(44, 30, seed=11)
    print(f"[case] fast-sediment {NAMES[form]} {sc.name}", flush=True)
    g, o = make_pair(sc, form=form, sediment=True, bed_change=True)
    g.step(8)
    o.step(8)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("Z", "U", "V"):
        d = np.max(np.abs(g.get(f) - o.get(f)))
        print(f"[{f}] max abs error {d:.3e}", flush=True)
        assert d <= TOL[form], (f, d)
    for f in ("FS", "DH"):
        a, b = g.get(f), o.get(f)
        scale = np.abs(b).max() + 1e-300
        print(f"[{f}] relative error {np.max(np.abs(a - b)) / scale:.3e}", flush=True)
        assert np.max(np.abs(a - b)) <= REL_TOL[form] * scale, (f, np.max(np.abs(a - b)), scale)


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
def test_fast_form_one_hour_config1(form):
    """Config 1, one simulated hour (1800 steps): the accuracy bar of the
    mixed-precision form (|dZ| <= 1 mm, |dU|, |dV| <= 1e-4 m/s)."""
    sc = S.config1()
    print(f"[case] 1h config1 {NAMES[form]}", flush=True)
    g, o = make_pair(sc, form=form)
    for _ in range(18):
        g.step(100)
        o.step(100)
    dz = np.max(np.abs(g.get("Z") - o.get("Z")))
    du = np.max(np.abs(g.get("U") - o.get("U")))
    dv = np.max(np.abs(g.get("V") - o.get("V")))
    print(f"[1h] {NAMES[form]} dZ={dz:.3e} dU={du:.3e} dV={dv:.3e}", flush=True)
    assert dz <= 1e-3 and du <= 1e-4 and dv <= 1e-4
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))


SOLVERS = [(engine.SOLVER_PARTITION, 2), (engine.SOLVER_PARTITION, 4), (engine.SOLVER_PARTITION, 8),
           (engine.SOLVER_CR, 0)]


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
@pytest.mark.parametrize("kind,param", SOLVERS, ids=["part2", "part4", "part8", "cr"])
def test_every_line_solver_close_to_oracle(form, kind, param):
    """Each line solver, on every sweep (hydraulic and FS), against the
    oracle at the fast forms' tolerance."""
    sc = S.meander_islands(160, 96, n_islands=8)
    sc.sediment_start = sc.bed_change_start = 4.0
    print(f"[case] solver {NAMES[form]} kind={kind} K/threads={param}", flush=True)
    g, o = make_pair(sc, form=form, sediment=True, bed_change=True)
    for d in (0, 1):
        for hyd in (0, 1):
            engine.set_solver(g, d, hyd, kind, param)
    g.step(10)
    o.step(10)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("Z", "U", "V"):
        d = np.max(np.abs(g.get(f) - o.get(f)))
        print(f"[{f}] max abs error {d:.3e}", flush=True)
        assert d <= TOL[form], (f, d)
    a, b = g.get("FS"), o.get("FS")
    scale = np.abs(b).max() + 1e-300
    print(f"[FS] relative error {np.max(np.abs(a - b)) / scale:.3e}", flush=True)
    assert np.max(np.abs(a - b)) <= REL_TOL[form] * scale
