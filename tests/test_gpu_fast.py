"""GPU fast forms against the CPU oracle.

HD_FORM_FP64_FAST solves the z-only (Schur) system of each line by cyclic
reduction in FP64; HD_FORM_FP32_REFINE solves it in FP32 and adds one FP64
residual-refinement solve.  Both differ from the reference's interleaved
Thomas only by round-off, so they are compared with tolerances:

  FP64_FAST:   |dZ|, |dU|, |dV| <= 1e-9   after the stated steps
  FP32_REFINE: |dZ| <= 1e-6 m, |dU|, |dV| <= 1e-6 m/s
  wet/dry mask and segment tables: identical
"""
import numpy as np
import pytest

from hydist_c_amd import scenarios as S
from hydist_c_amd.abi import HD_FORM_FP32_REFINE, HD_FORM_FP64_FAST
from parity_util import make_pair

pytestmark = pytest.mark.gpu

TOL = {HD_FORM_FP64_FAST: 1e-9, HD_FORM_FP32_REFINE: 1e-6}
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
        assert d <= TOL[form], (f, d)


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
def test_fast_form_wetting_drying(form):
    """Tidal flats (cells dry and re-wet): the upwind switches on
    H_moi <= H_TINH (UVZSolver_multithread.cpp:186-201, 252-268) flip
    under round-off near the front, so differences grow beyond round-off.
    Bar: |dZ| <= 1e-4 m, |dU|, |dV| <= 1e-4 m/s after 120 steps; mask
    differences are reported and must stay below 1% of the wet cells."""
    sc = S.tidal_flats(64, 40)
    print(f"[case] fast-flats {NAMES[form]}", flush=True)
    g, o = make_pair(sc, form=form)
    g.step(120)
    o.step(120)
    kg, ko = g.get("KHOUOT"), o.get("KHOUOT")
    nd = int((kg != ko).sum())
    print(f"[flats] {NAMES[form]} mask differences: {nd}", flush=True)
    assert nd <= 0.01 * (ko == 0).sum()
    for f in ("Z", "U", "V"):
        d = np.max(np.abs(g.get(f) - o.get(f)))
        print(f"[flats] {NAMES[form]} max|d{f}| = {d:.3e}", flush=True)
        assert d <= 1e-4, (f, d)


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
def test_fast_form_sediment(form):
    sc = S.open_box(44, 30, seed=11)
    print(f"[case] fast-sediment {NAMES[form]} {sc.name}", flush=True)
    g, o = make_pair(sc, form=form, sediment=True, bed_change=True)
    g.step(8)
    o.step(8)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("Z", "U", "V"):
        assert np.max(np.abs(g.get(f) - o.get(f))) <= TOL[form], f
    for f in ("FS", "DH"):
        a, b = g.get(f), o.get(f)
        scale = np.abs(b).max() + 1e-300
        assert np.max(np.abs(a - b)) <= 1e3 * TOL[form] * scale, (f, np.max(np.abs(a - b)), scale)


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
