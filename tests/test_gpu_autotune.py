"""The solver autotuner only picks workgroup sizes: any choice gives the
same bits, so tuning never changes a run's output."""
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
    a = engine.Simulation(sc, form=abi.HD_FORM_FP64_FAST, sediment=True, bed_change=True)
    b = engine.Simulation(sc, form=abi.HD_FORM_FP64_FAST, sediment=True, bed_change=True)
    a.init()
    b.init()
    rep = engine.autotune(b)
    assert len(rep["chosen"]) == 4 and all(t in (128, 256, 512) for t in rep["chosen"]), rep
    a.step(6)
    b.step(6)
    for f in ("U", "V", "Z", "FS"):
        assert bits_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()
