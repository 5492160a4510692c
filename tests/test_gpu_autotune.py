"""Line-solver choice.  Cyclic reduction's workgroup size changes no bit;
the partition solver's K and the choice between the two solvers change
results at rounding level only (within the fast forms' tolerance), and a
choice, once made, gives the same bits every time."""
import numpy as np
import pytest

from hydist_c_amd import abi, engine, scenarios as S
from parity_util import bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [abi.HD_FORM_FP64_FAST, abi.HD_FORM_FP32_REFINE])
def test_solver_workgroup_size_does_not_change_results(form):
    print(f"[case] autotune form={form}", flush=True)
    sc = S.open_box(96, 64, seed=11, n_islands=5)
    sc.sediment_start = sc.bed_change_start = 4.0
    runs = []
    for threads in (0, 64, 512):
        g = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
        g.init()
        for d in (0, 1):
            for hyd in (0, 1):
                engine.set_solver_threads(g, d, hyd, threads)
        g.step(8)
        runs.append({f: g.get(f) for f in ("U", "V", "Z", "FS")})
        g.close()
    for r in runs[1:]:
        for f in r:
            assert bits_equal(r[f], runs[0][f]), f


def test_autotune_reports_and_keeps_results():
    print("[case] autotune report", flush=True)
    sc = S.open_box(96, 64, seed=11, n_islands=5)
    sc.sediment_start = sc.bed_change_start = 4.0
    sims = [engine.Simulation(sc, form=abi.HD_FORM_FP64_FAST, sediment=True, bed_change=True) for _ in range(3)]
    a, b, c = sims
    for g in sims:
        g.init()
    rep = engine.autotune(b)
    chosen = rep["chosen"]
    cands = {"part/2", "part/4", "part/8", "cr/128", "cr/256", "cr/512", "seg/0"}
    assert len(chosen) == 4 and all(x in cands for x in chosen), rep
    # the same choices set by hand give the tuned run's bits
    kinds = {"cr": engine.SOLVER_CR, "part": engine.SOLVER_PARTITION, "seg": engine.SOLVER_SEGMENTS}
    for (d, hyd), x in zip(((1, 1), (0, 1), (1, 0), (0, 0)), chosen):
        kind, param = x.split("/")
        engine.set_solver(c, d, hyd, kinds[kind], int(param))
    for g in sims:
        g.step(6)
    for f in ("U", "V", "Z", "FS"):
        assert bits_equal(b.get(f), c.get(f)), f
        assert np.max(np.abs(a.get(f) - b.get(f))) <= 1e-9 * max(1.0, np.abs(a.get(f)).max()), f
    for g in sims:
        g.close()


def test_set_solver_rejects_bad_choices():
    print("[case] set_solver arguments", flush=True)
    sc = S.open_box(40, 28)
    g = engine.Simulation(sc, form=abi.HD_FORM_FP64_FAST)
    g.init()
    for d, hyd, kind, param in ((2, 1, engine.SOLVER_PARTITION, 4), (1, 1, engine.SOLVER_PARTITION, 3),
                                (1, 1, engine.SOLVER_CR, 96), (1, 0, engine.SOLVER_CR, 1024), (0, 1, 7, 0),
                                (0, 1, engine.SOLVER_DEFAULT, 4)):
        with pytest.raises(engine.HydistError):
            engine.set_solver(g, d, hyd, kind, param)
    engine.set_solver(g, 0, 1, engine.SOLVER_PARTITION, 8)
    engine.set_solver(g, 1, 1, engine.SOLVER_DEFAULT, 0)
    g.step(2)
    g.close()
