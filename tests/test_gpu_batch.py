"""Ensembles and whole-step graphs.

A batch runs n scenarios of one shape with one launch per stage; each
scenario's results must equal the same scenario run alone (bit for bit:
every line's arithmetic is independent of the batch).  A graph-captured
step must equal the eager launch sequence bit for bit, replay after
replay, including across the step where sediment and bed change switch
on (a new graph is captured for the new stage pattern)."""
import numpy as np
import pytest

from hydist_c_amd import abi, engine, scenarios as S
from parity_util import bits_equal

pytestmark = pytest.mark.gpu


FIELDS = ["U", "V", "Z", "FS", "DH", "KHOUOT"]
FORMS = {"plain": abi.HD_FORM_FP64_PLAIN, "fast": abi.HD_FORM_FP64_FAST, "fp32": abi.HD_FORM_FP32_REFINE}


def members(n, M=96, N=48, start=20.0):
    out = []
    for k in range(n):
        s = S.ensemble_member(k, M=M, N=N)
        s.sediment_start = s.bed_change_start = start
        out.append(s)
    return out


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.mark.parametrize("form", ["plain", "fast", "fp32"])
def test_batch_equals_single_runs(form):
    _case("batch vs single", form=form)
    scens = members(4)
    b = engine.Batch(scens, form=FORMS[form], sediment=True, bed_change=True)
    b.init()
    b.step(20)
    for k, sc in enumerate(scens):
        g = engine.Simulation(sc, form=FORMS[form], sediment=True, bed_change=True)
        g.init()
        g.step(20)
        for f in FIELDS:
            a = b.get(k, f)
            e = g.get(f).astype(np.float64)
            assert bits_equal(a, e), (form, k, f, float(np.nanmax(np.abs(a - e))))
        g.close()
    b.close()


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_graph_replays_equal_eager(form):
    _case("graph vs eager", form=form)
    scens = members(3)
    eager = engine.Batch(scens, form=FORMS[form], sediment=True, bed_change=True)
    graph = engine.Batch(scens, form=FORMS[form], sediment=True, bed_change=True, graph=True)
    eager.init()
    graph.init()
    # steps 1-9 before sediment, 10 switches it on, then several replays of
    # the sediment graph; every step is compared (copy-out synchronises)
    for step in range(16):
        eager.step()
        graph.step()
        assert eager.t == graph.t
        for k in range(len(scens)):
            for f in ("Z", "U", "FS"):
                assert bits_equal(eager.get(k, f), graph.get(k, f)), (form, step, k, f)
    eager.close()
    graph.close()


def test_batch_rejects_mixed_shapes():
    _case("batch shapes")
    a = S.ensemble_member(0, M=64, N=32)
    b = S.ensemble_member(1, M=64, N=40)
    with pytest.raises(engine.HydistError):
        engine.Batch([a, b])


def test_solver_change_recaptures_graph():
    """A solver change after graph steps takes effect in the next graph
    step: the graph run equals an eager run with the same solver sequence,
    and differs (at rounding level) from a run that kept the first
    solver."""
    _case("graph recapture after set_solver")
    sc = S.open_box(64, 40, seed=11, n_islands=3)
    runs = []
    for graph, change in ((True, True), (False, True), (False, False)):
        g = engine.Simulation(sc, form=abi.HD_FORM_FP64_FAST)
        if graph:
            engine.check(engine.lib().hd_set_graph(g.h, 1), g.h)
        g.init()
        g.step(3)
        if change:
            for d in (0, 1):
                engine.set_solver(g, d, 1, engine.SOLVER_CR, 256)
        g.step(4)
        runs.append({f: g.get(f) for f in ("Z", "U", "V")})
        g.close()
    graph_run, eager_run, kept = runs
    for f in graph_run:
        assert bits_equal(graph_run[f], eager_run[f]), f
    assert not all(bits_equal(eager_run[f], kept[f]) for f in kept), "the solvers round alike here"


@pytest.mark.parametrize("form", ["fast", "fp32"])
def test_256_members_on_a_small_mesh(form):
    """The 256-member ensemble's context shape (one launch per stage over
    256 scenarios, the scenario as the grid's outer dimension, whole-step
    graphs) on a small mesh: sampled members equal their single runs bit
    for bit."""
    _case("256 members", form=form)
    scens = members(256, M=24, N=16, start=4.0)
    b = engine.Batch(scens, form=FORMS[form], sediment=True, bed_change=True, graph=True)
    b.init()
    b.step(6)
    for k in (0, 1, 128, 255):
        g = engine.Simulation(scens[k], form=FORMS[form], sediment=True, bed_change=True)
        g.init()
        g.step(6)
        for f in FIELDS:
            assert bits_equal(b.get(k, f), g.get(f).astype(np.float64)), (form, k, f)
        g.close()
    b.close()
