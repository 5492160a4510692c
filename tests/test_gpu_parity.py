"""GPU parity: the gfx950 plain FP64 form against the CPU oracle.

Hydraulic fields are compared bit for bit (every value is formed from
+ - * / sqrt in the reference's order with contraction off).  Fields that
involve pow/log (init Kx1/Ky1/hi, sediment FS/Kx/Ky/dH) are compared with
stated tolerances; the pow() init fields are then copied into the oracle
so the trajectories start from the same state."""
import numpy as np
import pytest

import oracle_lib as ol
from hydist_c_amd import engine, scenarios as S
from hydist_c_amd.abi import HD_QUIRK_Q11_LITERAL
from parity_util import HYDRO, bits_equal, diff_report, make_pair, ulp_diff

pytestmark = pytest.mark.gpu


def _print_case(case, **kw):
    print(f"[case] {case} {kw}", flush=True)


@pytest.mark.parametrize("maker", [lambda: S.open_box(40, 28), lambda: S.config1(),
                                   lambda: S.meander_islands(96, 64, n_islands=6)])
def test_init_matches_oracle(maker):
    sc = maker()
    _print_case("init", name=sc.name)
    g, o = make_pair(sc, seed_init=False)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("HTAIZ", "HTAIZ_BD", "HTDU", "HTDV", "HMOI"):
        assert bits_equal(g.get(f), o.get(f)), f
    # pow() on the device vs glibc: a few ulp
    for f in ("KX1", "KY1"):
        a, b = g.get(f), o.get(f)
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin), f
        assert ulp_diff(a[fin], b[fin]).max() <= 4, f
    hg, ho = g.get_boundary("HI"), o.get_boundary("HI")
    assert np.allclose(hg, ho, rtol=1e-13, atol=0)
    assert np.allclose(g.get_hmax(), o.get_hmax(), rtol=1e-15, atol=0)
    for d in (0, 1):
        cg, fg, lg = g.segments(d)
        co, fo, lo = o.segments(d)
        assert np.array_equal(cg, co) and np.array_equal(fg, fo) and np.array_equal(lg, lo)


@pytest.mark.parametrize("maker,steps", [
    (lambda: S.open_box(40, 28), 12),
    (lambda: S.open_box(57, 33, seed=9, types=(0, 2, 0, 2)), 12),
    (lambda: S.config1(), 30),
    (lambda: S.meander_islands(96, 64, n_islands=6), 20),
    (lambda: S.closed_basin(48, 36, z0_amp=0.3), 25),
    (lambda: S.tidal_flats(64, 40), 200),
])
def test_plain_fp64_bitwise(maker, steps):
    sc = maker()
    _print_case("plain_bitwise", name=sc.name, steps=steps)
    g, o = make_pair(sc)
    for k in range(steps):
        g.step()
        o.step()
    rep = diff_report(g, o, HYDRO)
    bad = {k: v for k, v in rep.items() if not v["bitwise"]}
    assert not bad, bad
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for d in (0, 1):
        assert all(np.array_equal(a, b) for a, b in zip(g.segments(d), o.segments(d)))


def test_plain_fp64_quirk_q11_literal():
    sc = S.open_box(40, 28, types=(2, 0, 0, 0))
    _print_case("q11_literal", name=sc.name)
    g, o = make_pair(sc, quirks=HD_QUIRK_Q11_LITERAL)
    g.step(4)
    o.step(4)
    rep = diff_report(g, o, HYDRO)
    assert all(v["bitwise"] for v in rep.values()), rep


def test_half_steps_match_individually():
    """Stage parity: after each half-step separately."""
    sc = S.open_box(36, 26, seed=5)
    _print_case("halfsteps", name=sc.name)
    g, o = make_pair(sc)
    t = 0.0
    for k in range(3):
        t += 1.0
        g.halfstep_uz(t)
        o.halfstep_uz(t)
        rep = diff_report(g, o, HYDRO)
        assert all(v["bitwise"] for v in rep.values()), ("uz", k, rep)
        t += 1.0
        g.halfstep_vz(t)
        o.halfstep_vz(t)
        rep = diff_report(g, o, HYDRO)
        assert all(v["bitwise"] for v in rep.values()), ("vz", k, rep)


def test_sediment_and_bed_change():
    """Hydraulics stay bitwise with sediment on (nothing hydraulic reads FS);
    FS, Kx, Ky and dH agree to the stated tolerance (device log/pow)."""
    sc = S.open_box(44, 30, seed=11)
    _print_case("sediment", name=sc.name)
    g, o = make_pair(sc, sediment=True, bed_change=True)
    g.step(10)
    o.step(10)
    rep = diff_report(g, o, HYDRO + ["VTH"])
    assert all(v["bitwise"] for v in rep.values()), rep
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f, rtol in (("FS", 1e-10), ("KX", 1e-12), ("KY", 1e-12), ("DH", 1e-9), ("TFS", 1e-10)):
        a, b = g.get(f), o.get(f)
        scale = np.abs(b).max() + 1e-300
        assert np.max(np.abs(a - b)) <= rtol * scale, (f, np.max(np.abs(a - b)), scale)
    assert np.abs(o.get("DH")).max() > 0 and np.abs(o.get("FS")).max() > 0


def test_lake_at_rest_gpu():
    sc = S.closed_basin(40, 30)
    g, o = make_pair(sc)
    g.step(10)
    for f in ("Z", "U", "V"):
        assert not np.any(g.get(f)), f


def test_drying_threshold_is_inclusive():
    """Reset_states_vertical dries a wet cell when H_moi <= H_TINH
    (support_funcs.cpp:142-188), equality included.  A lake at rest keeps
    z exactly 0, so H_moi is exactly the cell-centre depth (4 m here);
    with H_TINH set to that depth every interior wet cell sits exactly on
    the threshold and must dry, as in the oracle."""
    sc = S.closed_basin(30, 22, depth=4.0)
    sc.coeff_overrides = dict(sc.coeff_overrides, H_TINH=4.0)
    g, o = make_pair(sc)
    wet0 = int(np.sum(g.get("KHOUOT") == 0))
    g.step(2)
    o.step(2)
    kh = g.get("KHOUOT")
    assert np.array_equal(kh, o.get("KHOUOT"))
    assert wet0 > 0 and int(np.sum(kh == 0)) < wet0 // 2, (wet0, int(np.sum(kh == 0)))
    for f in HYDRO:
        assert bits_equal(g.get(f), o.get(f)), f


def test_plain_fp64_one_hour_config1_bitwise():
    """Config 1 for one simulated hour (1800 steps): still bit for bit."""
    sc = S.config1()
    _print_case("plain_1h", name=sc.name)
    g, o = make_pair(sc)
    for _ in range(18):
        g.step(100)
        o.step(100)
    rep = diff_report(g, o, HYDRO)
    assert all(v["bitwise"] for v in rep.values()), rep
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))


def test_mass_balance_gpu():
    sc = S.closed_basin(48, 36, z0_amp=0.2)
    g, o = make_pair(sc)
    v0, _ = g.mass_balance()
    g.step(20)
    o.step(20)
    v1, q = g.mass_balance()
    vo, qo = o.mass_balance()
    assert q == 0.0
    assert abs(v1 - v0) < 1e-9 * v0
    assert v1 == vo  # same state, same summation order


@pytest.mark.parametrize("mode", ["kenhhepd", "kenhhepng"])
def test_narrow_channel_modes_bitwise(mode):
    """Channel mode (engine.cpp:24-30,77-80,190-193): no boundary update,
    and the narrowed sweep range of the chosen direction."""
    _print_case("channel", mode=mode)
    sc = S.meander_channel(96, 48, seed=7, amplitude=6.0, wavelength=96.0, half_width=12.0)
    sc.options = {"channel": True, mode: True}
    g, o = make_pair(sc)
    g.step(15)
    o.step(15)
    rep = diff_report(g, o, HYDRO)
    assert all(v["bitwise"] for v in rep.values()), rep


def test_stage_entry_points_match_oracle():
    """hd_sediment_j / hd_sediment_i / hd_bedload against the oracle's
    stage blocks on the same mid-run state (FS and dH involve pow/log, so a
    relative tolerance; the hydraulic fields are untouched)."""
    _print_case("stages")
    sc = S.open_box(40, 28, seed=3)
    sc.sediment_start = sc.bed_change_start = 1e9  # stages run only when called
    g, o = make_pair(sc, sediment=True, bed_change=True)
    g.step(8)
    o.step(8)
    for name, t in (("sediment_j", 17.0), ("sediment_i", 18.0), ("bedload", 18.0)):
        g.stage(name, t)
        o.stage(name, t)
    for f in ("FS", "DH", "QBX", "QBY", "VTH"):
        a, b = g.get(f), o.get(f)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-18), (f, float(np.max(np.abs(a - b))))
    assert float(np.abs(g.get("FS")).max()) > 0
    rep = diff_report(g, o, HYDRO)
    assert all(v["bitwise"] for v in rep.values()), rep


def test_run_with_output_callback():
    _print_case("run_cb")
    sc = S.open_box(40, 28, seed=3)
    g = engine.Simulation(sc)
    g.init()
    seen = []
    assert g.run(20.0, interval=4, callback=lambda t: seen.append(t) or False)
    assert seen == [4.0, 8.0, 12.0, 16.0, 20.0] and g.t == 20.0
    stopped = []
    assert not g.run(40.0, interval=2, callback=lambda t: stopped.append(t) or t >= 26.0)
    assert stopped[-1] == 26.0 and g.t == 26.0
    g.close()
