"""The benchmark configurations at their full sizes, on the GPU.

C2  2048 x 1024 meander + 30 islands, sediment and FP64 bed change from
    t = 300 s, one simulated hour (1800 steps of dT = 2 s), every form
    against the oracle.  Masks and segment tables identical; the plain
    form's U, V, Z, H_moi bit for bit; the fast forms within the accuracy
    bar (|dZ| <= 1 mm, |dU|, |dV| <= 1e-4 m/s) and within the tighter
    per-form bars below; FS, bed change and the domain volume (mass
    balance) within stated relative bars.
C3  4096 x 4096 braided river, suspended sediment on, dT = 1 s, 600
    steps: the plain form against the oracle over the first 100 steps
    (bit for bit on the hydraulics; the oracle takes about a second per
    step at this size), FP32 + refine against the oracle there and against
    the plain form after all 600 steps, errors of U, V, Z and FS reported.
C4  256 scenarios of 512 x 256 in one batched context (one launch per
    stage over all of them, one graph launch per step), fast FP64 and
    FP32 + refine, one simulated hour: 8 sampled members equal single
    runs of the same scenario bit for bit.
C5  what-if requests drawn from the serving distribution (sides 128-2048,
    sediment for half, precision picked by the requested tolerance), run
    through the serving loop; durations are capped at 300 simulated
    seconds to bound the test: every completed run equals the same run
    alone, bit for bit.
"""
import threading
import time

import numpy as np
import pytest

import oracle_lib as ol
from hydist_c_amd import abi, engine, scenarios as S
from parity_util import bits_equal, make_pair

pytestmark = pytest.mark.gpu

PLAIN, FAST, FP32 = abi.HD_FORM_FP64_PLAIN, abi.HD_FORM_FP64_FAST, abi.HD_FORM_FP32_REFINE
NAMES = {PLAIN: "plain", FAST: "fp64_fast", FP32: "fp32_refine"}

# Bars after one simulated hour of C2, about 100 times what is observed:
# hydraulics (m, m/s) fast 1.2e-15, FP32 + refine 3.7e-12; FS relative to
# its largest value 3e-16 (plain), 8.6e-16 (fast), 1.0e-12 (FP32 + refine);
# bed change relative 2.4e-16, 5.9e-16, 3.1e-13; domain volume relative 0,
# 0, 1.3e-13.
HYD_TOL = {FAST: 1e-13, FP32: 5e-10}
FS_REL = {PLAIN: 1e-13, FAST: 1e-13, FP32: 1e-10}
DH_REL = {PLAIN: 1e-13, FAST: 1e-13, FP32: 5e-11}
# domain volume: the plain form's equals the oracle's bit for bit (same
# state, same summation order)
VOL_REL = {PLAIN: 0.0, FAST: 1e-13, FP32: 1e-11}


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


class Background:
    """Runs fn() in a thread (the oracle's OpenMP work releases the GIL, so
    GPU runs proceed meanwhile)."""

    def __init__(self, fn):
        self.out, self.err = None, None
        self.t = threading.Thread(target=self._run, args=(fn,), daemon=True)
        self.t.start()

    def _run(self, fn):
        try:
            self.out = fn()
        except BaseException as e:  # reported in result()
            self.err = e

    def result(self):
        self.t.join()
        if self.err is not None:
            raise self.err
        return self.out


def state(sim, names=("Z", "U", "V", "HMOI", "FS", "DH", "KHOUOT")):
    out = {f: sim.get(f) for f in names}
    out["segments"] = [sim.segments(d) for d in (0, 1)]
    out["volume"] = sim.mass_balance()[0]
    return out


def absdiff(a, b):
    """Largest |a - b| over the entries finite in both; NaNs (FS and the
    bed change on dry cells next to the wet ones) must sit at the same
    places."""
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN positions differ"
    fin = np.isfinite(a) & np.isfinite(b)
    return float(np.max(np.abs(a[fin] - b[fin]))) if fin.any() else 0.0


def rel(a, b):
    """absdiff relative to b's largest finite magnitude."""
    d = absdiff(a, b)
    fin = np.isfinite(b)
    scale = float(np.max(np.abs(b[fin]))) if fin.any() else 0.0
    return d / scale if scale > 0 else d


def same_segments(a, b):
    return all(np.array_equal(x, y) for sa, sb in zip(a, b) for x, y in zip(sa, sb))


# ---------------------------------------------------------------------------
# C2
# ---------------------------------------------------------------------------
C2_STEPS = 1800


@pytest.fixture(scope="module")
def c2():
    sc = S.config2()
    g, o = make_pair(sc, form=PLAIN, sediment=True, bed_change=True)
    g.close()

    def run():
        for k in range(C2_STEPS // 300):
            o.step(300)
            print(f"[c2 oracle] step {(k + 1) * 300}", flush=True)
        return state(o)

    return sc, Background(run)


@pytest.mark.parametrize("form", [PLAIN, FAST, FP32], ids=NAMES.get)
def test_c2_one_hour_against_oracle(c2, form):
    sc, oracle = c2
    _case("C2 1h", form=NAMES[form], steps=C2_STEPS)
    g = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
    g.init()
    t0 = time.time()
    for k in range(C2_STEPS // 300):
        g.step(300)
        g.sync()
        print(f"[c2 {NAMES[form]}] step {(k + 1) * 300} ({time.time() - t0:.1f} s)", flush=True)
    got = state(g)
    g.close()
    ref = oracle.result()
    assert np.array_equal(got["KHOUOT"], ref["KHOUOT"])
    assert same_segments(got["segments"], ref["segments"])
    d = {f: absdiff(got[f], ref[f]) for f in ("Z", "U", "V", "HMOI")}
    fs, dh = rel(got["FS"], ref["FS"]), rel(got["DH"], ref["DH"])
    vol = abs(got["volume"] - ref["volume"]) / ref["volume"]
    print(f"[c2 {NAMES[form]}] max|dZ| {d['Z']:.3e} |dU| {d['U']:.3e} |dV| {d['V']:.3e} |dH_moi| {d['HMOI']:.3e} "
          f"FS rel {fs:.3e} dH rel {dh:.3e} (max|dH| {np.nanmax(np.abs(ref['DH'])):.3e}) volume rel {vol:.3e} "
          f"(volume {ref['volume']:.6e} m^3)", flush=True)
    if form == PLAIN:
        for f in ("Z", "U", "V", "HMOI"):
            assert bits_equal(got[f], ref[f]), f
    else:
        assert d["Z"] <= 1e-3 and d["U"] <= 1e-4 and d["V"] <= 1e-4, d
        assert max(d.values()) <= HYD_TOL[form], d
    assert fs <= FS_REL[form] and dh <= DH_REL[form], (fs, dh)
    assert vol <= VOL_REL[form], vol


# ---------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------
C3_STEPS, C3_ORACLE_STEPS = 600, 100


def test_c3_braided_sediment():
    sc = S.config3(sediment=True)
    _case("C3", steps=C3_STEPS, oracle_steps=C3_ORACLE_STEPS)
    g, o = make_pair(sc, form=PLAIN, sediment=True)

    def run():
        for k in range(C3_ORACLE_STEPS // 25):
            o.step(25)
            print(f"[c3 oracle] step {(k + 1) * 25}", flush=True)
        return state(o, ("Z", "U", "V", "HMOI", "FS", "KHOUOT"))

    oracle = Background(run)
    f = engine.Simulation(sc, form=FP32, sediment=True)
    f.init()
    for sim in (g, f):
        sim.step(C3_ORACLE_STEPS)
    at100 = {"plain": state(g, ("Z", "U", "V", "HMOI", "FS", "KHOUOT")),
             "fp32": state(f, ("Z", "U", "V", "HMOI", "FS", "KHOUOT"))}
    t0 = time.time()
    for k in range((C3_STEPS - C3_ORACLE_STEPS) // 100):
        g.step(100)
        f.step(100)
        g.sync()
        f.sync()
        print(f"[c3 gpu] step {C3_ORACLE_STEPS + (k + 1) * 100} ({time.time() - t0:.1f} s)", flush=True)
    end_plain = {n: g.get(n) for n in ("Z", "U", "V", "FS", "KHOUOT")}
    end_fp32 = {n: f.get(n) for n in ("Z", "U", "V", "FS", "KHOUOT")}
    g.close()
    f.close()
    ref = oracle.result()
    # plain form against the oracle over the first 100 steps
    assert np.array_equal(at100["plain"]["KHOUOT"], ref["KHOUOT"])
    assert same_segments(at100["plain"]["segments"], ref["segments"])
    for n in ("Z", "U", "V", "HMOI"):
        assert bits_equal(at100["plain"][n], ref[n]), n
    fs_plain = rel(at100["plain"]["FS"], ref["FS"])
    # FP32 + refine against the oracle (100 steps) and the plain form (600)
    e100 = {n: absdiff(at100["fp32"][n], ref[n]) for n in ("Z", "U", "V")}
    e100["FS_rel"] = rel(at100["fp32"]["FS"], ref["FS"])
    e600 = {n: absdiff(end_fp32[n], end_plain[n]) for n in ("Z", "U", "V")}
    e600["FS_rel"] = rel(end_fp32["FS"], end_plain["FS"])
    print(f"[c3] plain vs oracle @{C3_ORACLE_STEPS}: hydraulics bitwise, FS rel {fs_plain:.3e}; "
          f"fp32 vs oracle @{C3_ORACLE_STEPS} {e100}; fp32 vs plain @{C3_STEPS} {e600}; "
          f"max FS {np.nanmax(np.abs(end_plain['FS'])):.3e}", flush=True)
    assert fs_plain <= FS_REL[PLAIN]
    assert np.array_equal(at100["fp32"]["KHOUOT"], ref["KHOUOT"])
    assert np.array_equal(end_fp32["KHOUOT"], end_plain["KHOUOT"])
    for e in (e100, e600):
        assert e["Z"] <= HYD_TOL[FP32] and e["U"] <= HYD_TOL[FP32] and e["V"] <= HYD_TOL[FP32], e
        assert e["FS_rel"] <= FS_REL[FP32], e
    assert np.nanmax(np.abs(end_plain["FS"])) > 0, "no suspended sediment in the case"


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------
C4_SCENARIOS, C4_STEPS = 256, 1800
C4_SAMPLE = (0, 1, 37, 100, 128, 191, 222, 255)


@pytest.mark.parametrize("form", [FAST, FP32], ids=NAMES.get)
def test_c4_ensemble_256_members_equal_single_runs(form):
    scens = [S.ensemble_member(k) for k in range(C4_SCENARIOS)]
    _case("C4", form=NAMES[form], scenarios=C4_SCENARIOS, mesh=(scens[0].M, scens[0].N), steps=C4_STEPS)
    b = engine.Batch(scens, form=form, graph=True)
    b.init()
    t0 = time.time()
    for k in range(C4_STEPS // 300):
        b.step(300)
        b.sync()
        print(f"[c4 {NAMES[form]}] step {(k + 1) * 300} ({time.time() - t0:.1f} s)", flush=True)
    got = {k: {f: b.get(k, f) for f in ("Z", "U", "V", "KHOUOT")} for k in C4_SAMPLE}
    b.close()
    for k in C4_SAMPLE:
        g = engine.Simulation(scens[k], form=form)
        g.init()
        g.step(C4_STEPS)
        for f in ("Z", "U", "V", "KHOUOT"):
            assert bits_equal(got[k][f], g.get(f).astype(np.float64)), (k, f)
        assert np.isfinite(got[k]["Z"]).all()
        g.close()
    print(f"[c4 {NAMES[form]}] {len(C4_SAMPLE)} sampled members equal their single runs", flush=True)


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
C5_REQUESTS, C5_MAX_T = 12, 300.0


def test_c5_serving_requests_equal_alone():
    reqs = []
    for k in range(C5_REQUESTS):
        sc, T, form, sed = S.serving_request(k, fp64_form=1)
        reqs.append((sc, min(T, C5_MAX_T), form, sed))
    _case("C5", requests=[(sc.M, sc.N, T, NAMES[f], sed) for sc, T, f, sed in reqs])
    assert {f for _, _, f, _ in reqs} == {FAST, FP32}
    assert {s for *_, s in reqs} == {True, False}
    srv = engine.Server(hbm_budget=64 << 30, max_active=8, steps_per_quantum=8, use_graph=False)
    rng = np.random.default_rng(7)
    ids = []
    for sc, T, f, sed in reqs:
        time.sleep(float(rng.exponential(1.0 / 20.0)))  # 20 requests/s
        ids.append(srv.submit(sc, T, form=f, sediment=sed, bed_change=sed))
    for rid in ids:
        st = srv.wait(rid, timeout=240.0)
        assert st == engine.RUN_DONE, (rid, st, srv.error(rid))
    stats = srv.stats()
    print(f"[c5] {stats}", flush=True)
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    for rid, (sc, T, f, sed) in zip(ids, reqs):
        g = engine.Simulation(sc, form=f, sediment=sed, bed_change=sed, Tmax=T)
        g.init()
        while g.t < int(T):
            g.step()
        for name in ("U", "V", "Z", "FS", "DH"):
            assert bits_equal(srv.result(rid, name), g.get(name)), (rid, sc.name, name)
        g.close()
    srv.close()
