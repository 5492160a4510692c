"""Randomised parity: scenarios drawn from a seeded distribution of mesh
sizes (odd and even, about 20-180 cells a side), island counts, depths,
boundary types on each side (discharge or water level), inflows and
sediment switches, each run for a few steps through the C ABI and through
the oracle.  The plain form must match bit for bit (U, V, Z, H_moi, the
wet/dry mask and the segment tables; FS and dH, which pass through
pow/log, within 1e-14 of their largest values); the fast forms within
the bars below, about 100 times the largest error seen over the draw
(fast FP64: 1.8e-15 on the hydraulics, 2.9e-14 relative on FS and dH;
FP32 + refine: 3.0e-14 and 5.4e-13)."""
import numpy as np
import pytest

from hydist_c_amd import abi, scenarios as S
from parity_util import HYDRO, bits_equal, make_pair

pytestmark = pytest.mark.gpu

PLAIN, FAST, FP32 = abi.HD_FORM_FP64_PLAIN, abi.HD_FORM_FP64_FAST, abi.HD_FORM_FP32_REFINE
NAMES = {PLAIN: "plain", FAST: "fp64_fast", FP32: "fp32_refine"}
HYD_TOL = {FAST: 2e-13, FP32: 3e-12}
REL_TOL = {PLAIN: 1e-14, FAST: 3e-12, FP32: 5e-11}
STEPS = 8


def random_scenario(k):
    """Draw k: (scenario, sediment and bed change on)."""
    rng = np.random.default_rng(90_000 + k)
    M, N = int(rng.integers(20, 91)), int(rng.integers(16, 71))
    kind = k % 3
    if kind == 0:
        types = tuple(int(rng.choice([0, 2])) for _ in range(4))
        sc = S.open_box(M, N, seed=k, n_islands=int(rng.integers(0, 6)),
                        depth=(float(rng.uniform(1.0, 3.0)), float(rng.uniform(4.0, 8.0))), types=types,
                        Q_up=-float(rng.uniform(50, 300)), Q_left=-float(rng.uniform(20, 150)),
                        name=f"random{k}_box_{M}x{N}_{''.join(map(str, types))}")
    elif kind == 1:
        M2 = 2 * M
        sc = S.meander_islands(M2, N, seed=k, n_islands=int(rng.integers(1, 8)), Q=-4000.0 * N / 1024,
                               sediment_start=0.0, name=f"random{k}_islands_{M2}x{N}")
    else:
        sc = S.tidal_flats(M + 20, N + 10, seed=k, name=f"random{k}_flats_{M + 20}x{N + 10}")
    sc.sediment_start = sc.bed_change_start = 0.0
    return sc, bool(rng.random() < 0.7)


DRAWS = list(range(12))


def rel(a, b):
    fin = np.isfinite(a) & np.isfinite(b)
    scale = max(float(np.max(np.abs(b[fin]))) if fin.any() else 0.0, 1e-300)
    return float(np.max(np.abs(a[fin] - b[fin]))) / scale if fin.any() else 0.0


@pytest.mark.parametrize("form", [PLAIN, FAST, FP32], ids=NAMES.get)
@pytest.mark.parametrize("k", DRAWS)
def test_random_scenario_against_oracle(k, form):
    sc, sed = random_scenario(k)
    print(f"[case] random draw {k} {sc.name} sediment={sed} form={NAMES[form]}", flush=True)
    g, o = make_pair(sc, form=form, sediment=sed, bed_change=sed)
    g.step(STEPS)
    o.step(STEPS)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for d in (0, 1):
        for a, b in zip(g.segments(d), o.segments(d)):
            assert np.array_equal(np.asarray(a), np.asarray(b))
    for f in HYDRO:
        a, b = g.get(f), o.get(f)
        assert np.array_equal(np.isnan(a), np.isnan(b)), f
        if form == PLAIN:
            assert bits_equal(a, b), f
        else:
            fin = np.isfinite(a)
            err = float(np.max(np.abs(a[fin] - b[fin]))) if fin.any() else 0.0
            print(f"[{f}] max abs error {err:.3e}", flush=True)
            assert err <= HYD_TOL[form], (f, err)
    if sed:
        for f in ("FS", "DH"):
            e = rel(g.get(f), o.get(f))
            print(f"[{f}] relative error {e:.3e}", flush=True)
            assert e <= REL_TOL[form], (f, e)
    g.close()
