"""The serving loop: runs arrive while others are on the GPU, are admitted
by HBM footprint and slot count, advance together, and each result equals
the same run alone (bit for bit: a run's arithmetic never depends on what
else runs).  Admission failures and cancellation end a run cleanly."""
import time

import numpy as np
import pytest

from hydist_c_amd import abi, engine, scenarios as S
from parity_util import bits_equal

pytestmark = pytest.mark.gpu

PLAIN, FAST, FP32 = abi.HD_FORM_FP64_PLAIN, abi.HD_FORM_FP64_FAST, abi.HD_FORM_FP32_REFINE
FIELDS = ["U", "V", "Z", "FS", "DH", "KHOUOT"]


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def requests():
    """(scenario, Tmax, form, sediment) of a small mixed workload."""
    out = []
    a = S.open_box(40, 28, seed=3)
    a.sediment_start = a.bed_change_start = 10.0
    out.append((a, 40.0, PLAIN, True))
    out.append((S.meander_channel(96, 48, seed=7, amplitude=8.0, wavelength=96.0, half_width=12.0), 30.0, FAST, False))
    out.append((S.open_box(64, 40, seed=11, n_islands=5), 24.0, FP32, False))
    b = S.meander_channel(128, 64, seed=8, amplitude=10.0, wavelength=128.0, half_width=16.0)
    b.sediment_start = b.bed_change_start = 6.0
    out.append((b, 20.0, FAST, True))
    out.append((S.tidal_flats(64, 40, seed=6), 36.0, PLAIN, False))
    out.append((S.open_box(48, 48, seed=13), 16.0, PLAIN, False))
    return out


def alone(sc, T, form, sed):
    g = engine.Simulation(sc, form=form, sediment=sed, bed_change=sed, Tmax=T)
    g.init()
    while g.t < int(T):
        g.step()
    res = {f: g.get(f).astype(np.float64) for f in FIELDS}
    g.close()
    return res


def test_server_mix_equals_alone_runs():
    reqs = requests()
    fps = sorted(engine.footprint(sc) for sc, *_ in reqs)
    budget = fps[-1] + fps[-2] + 1  # at most two of the largest at once
    _case("server mix", n=len(reqs), budget=budget)
    srv = engine.Server(hbm_budget=budget, max_active=3, steps_per_quantum=4, use_graph=False)
    ids = [srv.submit(sc, T, form=f, sediment=s, bed_change=s) for sc, T, f, s in reqs[:3]]
    # the rest arrive while the first are on the GPU
    t0 = time.time()
    while srv.poll(ids[0])[0] == engine.RUN_QUEUED and time.time() - t0 < 60:
        time.sleep(0.001)
    ids += [srv.submit(sc, T, form=f, sediment=s, bed_change=s) for sc, T, f, s in reqs[3:]]
    for rid in ids:
        st = srv.wait(rid, timeout=120.0)
        assert st == engine.RUN_DONE, (rid, st, srv.error(rid))
    stats = srv.stats()
    assert stats["completed"] == len(reqs) and stats["failed"] == 0, stats
    assert 0 < stats["p50_s"] <= stats["p99_s"], stats
    expect_cells = sum(sc.M * sc.N * int(np.ceil(int(T) / 2.0)) for sc, T, *_ in reqs)
    assert stats["cell_steps"] == pytest.approx(expect_cells), stats
    for rid, (sc, T, f, s) in zip(ids, reqs):
        ref = alone(sc, T, f, s)
        for name in FIELDS:
            assert bits_equal(srv.result(rid, name), ref[name]), (rid, sc.name, name)
    srv.close()


def test_server_budget_and_cancel():
    a = S.open_box(40, 28, seed=3)
    big = S.open_box(64, 40, seed=11)
    _case("server budget/cancel")
    srv = engine.Server(hbm_budget=engine.footprint(a) + 1, max_active=1, steps_per_quantum=2, use_graph=False)
    r_big = srv.submit(big, 20.0)
    r_a = srv.submit(a, 200.0)
    r_c = srv.submit(a, 20.0)
    assert srv.wait(r_big, timeout=60.0) == engine.RUN_FAILED
    assert "budget" in srv.error(r_big)
    srv.cancel(r_c)
    assert srv.wait(r_c, timeout=60.0) == engine.RUN_CANCELLED
    srv.cancel(r_a)
    assert srv.wait(r_a, timeout=60.0) in (engine.RUN_CANCELLED, engine.RUN_DONE)
    with pytest.raises(engine.HydistError):
        srv.result(r_c, "Z")
    stats = srv.stats()
    assert stats["failed"] == 1 and stats["cancelled"] >= 1, stats
    srv.close()


def test_server_batches_runs_that_share_launches():
    """Runs of one mesh size, duration, form and options (depths, islands
    and hydrographs differ) that wait together are admitted as one batched
    context: every launch serves all of them, and each result equals the
    same run alone, bit for bit.  A larger run holds the budget first, so
    the four small ones are all queued when it leaves."""
    big = S.open_box(128, 96, seed=5)
    small = [S.open_box(40, 28, seed=20 + k) for k in range(4)]
    fp_big, fp_group = engine.footprint(big), engine.footprint(small[0], nscen=4)
    assert fp_group <= fp_big  # the four fit together once the large run has left
    _case("server batching", fp_big=fp_big, fp_group=fp_group)
    srv = engine.Server(hbm_budget=fp_big + 1, max_active=8, steps_per_quantum=4, use_graph=False)
    r_big = srv.submit(big, 10.0, form=FAST)
    ids = [srv.submit(sc, 16.0, form=FAST, sediment=True, bed_change=True) for sc in small]
    for rid in [r_big] + ids:
        st = srv.wait(rid, timeout=120.0)
        assert st == engine.RUN_DONE, (rid, st, srv.error(rid))
    stats = srv.stats()
    assert stats["batched"] == 4 and stats["completed"] == 5 and stats["failed"] == 0, stats
    for rid, sc in zip(ids, small):
        ref = alone(sc, 16.0, FAST, True)
        for name in FIELDS:
            assert bits_equal(srv.result(rid, name), ref[name]), (rid, sc.name, name)
    srv.close()


@pytest.mark.parametrize("nscen", [1, 3])
def test_footprint_is_what_a_context_allocates(nscen):
    """hd_footprint, which admission charges before anything is allocated,
    equals the device bytes the context then holds (after init and a step,
    with sediment), alone and batched."""
    _case("footprint", nscen=nscen)
    sc = S.meander_islands(72, 40, n_islands=3)
    sc.sediment_start = sc.bed_change_start = 2.0
    L = engine._server_lib()
    if nscen == 1:
        g = engine.Simulation(sc, form=FAST, sediment=True, bed_change=True, Tmax=8.0)
    else:
        g = engine.Batch([sc] * nscen, form=FAST, sediment=True, bed_change=True, Tmax=8.0)
    g.init()
    g.step(4)
    held = engine.C.c_size_t()
    engine.check(L.hd_memory_bytes(g.h, engine.C.byref(held)), g.h)
    assert held.value == engine.footprint(sc, nscen=nscen)
    g.close()
