"""The plain form's Thomas with one lane per system (HD_SOLVER_SEGMENTS)
against one lane per line and against the oracle: the same bits on every
field (the systems of a line are independent; the line walk carries
values only across rows no system reads), on meshes with many systems
per line, with one- and two-cell systems, drying fronts, every boundary
type and the sediment sweeps.  The tuner picks between the two by time
alone, so a tuned run equals an untuned one bit for bit.  In the fast FP64
form the same Thomas is one more solver choice, held to the oracle at the
fast form's bars."""
import numpy as np
import pytest

from hydist_c_amd import abi, engine, scenarios as S
from parity_util import HYDRO, bits_equal, make_pair
from test_gpu_edge_meshes import MESHES

pytestmark = pytest.mark.gpu

FIELDS = HYDRO + ["KHOUOT", "FS", "DH"]
CASES = {
    "box": (lambda: S.open_box(40, 28), 12),
    "box_z_q": (lambda: S.open_box(57, 33, seed=9, types=(0, 2, 0, 2)), 12),
    "flats": (lambda: S.tidal_flats(64, 40, seed=6), 20),
    "islands": (lambda: S.meander_islands(160, 96, n_islands=8), 8),
    "speckled35": (MESHES["speckled35"], 6),
    "stripes": (MESHES["stripes"], 6),
    "onecell": (MESHES["onecell"], 6),
}


def per_system(sim):
    for d in (0, 1):
        for hyd in (0, 1):
            engine.set_solver(sim, d, hyd, engine.SOLVER_SEGMENTS)


@pytest.mark.parametrize("name", list(CASES))
def test_thomas_per_system_matches_per_line_and_oracle(name):
    maker, steps = CASES[name]
    sc = maker()
    sc.sediment_start = sc.bed_change_start = 0.0
    print(f"[case] thomas per system {sc.name} steps={steps}", flush=True)
    g, o = make_pair(sc, sediment=True, bed_change=True)
    per_system(g)
    line = engine.Simulation(sc, sediment=True, bed_change=True)
    line.init()
    for sim in (g, o, line):
        sim.step(steps)
    for f in HYDRO + ["KHOUOT"]:
        assert bits_equal(g.get(f), o.get(f)), f
    for f in FIELDS:
        assert bits_equal(g.get(f), line.get(f)), f
    g.close()
    line.close()


def test_plain_tuning_changes_no_bit():
    sc = S.meander_islands(160, 96, n_islands=8)
    sc.sediment_start = sc.bed_change_start = 0.0
    print("[case] plain autotune", flush=True)
    a = engine.Simulation(sc, sediment=True, bed_change=True)
    b = engine.Simulation(sc, sediment=True, bed_change=True)
    a.init()
    b.init()
    a.step(2)
    b.step(2)
    rep = engine.autotune(a)
    assert any(k.endswith("thomas_seg") for k in rep) and any(k.endswith("thomas_line") for k in rep), rep
    assert all(c in ("default/0", "seg/0") for c in rep["chosen"]), rep
    a.step(6)
    b.step(6)
    for f in FIELDS:
        assert bits_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_per_system_thomas_is_for_the_fp64_forms():
    sim = engine.Simulation(S.open_box(24, 18), form=abi.HD_FORM_FP32_REFINE)
    with pytest.raises(engine.HydistError):
        engine.set_solver(sim, 0, 1, engine.SOLVER_SEGMENTS)
    sim.close()


def test_thomas_per_system_in_the_fast_form():
    """In the fast FP64 form the per-system Thomas is one more candidate
    of the tuner: the same FP64 Thomas on the fast form's assembled rows
    (the row sweeps then assemble and extract in kernels of their own),
    within the fast form's bars of the oracle."""
    sc = S.meander_islands(160, 96, n_islands=8)
    sc.sediment_start = sc.bed_change_start = 4.0
    print("[case] thomas per system, fast FP64", flush=True)
    g, o = make_pair(sc, form=abi.HD_FORM_FP64_FAST, sediment=True, bed_change=True)
    per_system(g)
    g.step(10)
    o.step(10)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("Z", "U", "V"):
        d = float(np.max(np.abs(g.get(f) - o.get(f))))
        print(f"[{f}] max abs error {d:.3e}", flush=True)
        assert d <= 1e-13, (f, d)
    a, b = g.get("FS"), o.get("FS")
    e = float(np.max(np.abs(a - b))) / (np.abs(b).max() + 1e-300)
    print(f"[FS] relative error {e:.3e}", flush=True)
    assert e <= 1e-12
    rep = engine.autotune(g)
    assert any(k.endswith("/seg0") for k in rep), rep
    g.close()


def test_thomas_per_system_in_a_batch_with_graphs():
    """Batched scenarios (the list is made per scenario) under whole-step
    graphs: every member equals its single per-line run bit for bit."""
    scens = []
    for k in range(3):
        s = S.ensemble_member(k, M=96, N=48)
        s.sediment_start = s.bed_change_start = 10.0
        scens.append(s)
    print("[case] thomas per system, batch of 3 with graphs", flush=True)
    b = engine.Batch(scens, sediment=True, bed_change=True, graph=True)
    per_system(b)
    b.init()
    b.step(12)
    for k, sc in enumerate(scens):
        g = engine.Simulation(sc, sediment=True, bed_change=True)
        g.init()
        g.step(12)
        for f in ["U", "V", "Z", "FS", "DH", "KHOUOT"]:
            assert bits_equal(b.get(k, f), g.get(f).astype(np.float64)), (k, f)
        g.close()
    b.close()


def test_fast_form_thomas_per_system_in_a_batch_with_graphs():
    """The fast FP64 form's per-system Thomas in a batch under whole-step
    graphs: every member equals its single run with the same solver, bit
    for bit (a scenario's arithmetic does not depend on its batch)."""
    scens = []
    for k in range(3):
        s = S.ensemble_member(k, M=96, N=48)
        s.sediment_start = s.bed_change_start = 10.0
        scens.append(s)
    fast = abi.HD_FORM_FP64_FAST
    print("[case] thomas per system, fast FP64, batch of 3 with graphs", flush=True)
    b = engine.Batch(scens, form=fast, sediment=True, bed_change=True, graph=True)
    per_system(b)
    b.init()
    b.step(12)
    for k, sc in enumerate(scens):
        g = engine.Simulation(sc, form=fast, sediment=True, bed_change=True)
        per_system(g)
        g.init()
        g.step(12)
        for f in ["U", "V", "Z", "FS", "DH", "KHOUOT"]:
            assert bits_equal(b.get(k, f), g.get(f).astype(np.float64)), (k, f)
        g.close()
    b.close()
