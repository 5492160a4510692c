"""Closed-form checks that pin single kernel restatements of the oracle
(smoothing, Find_VTH, hesoK, Calculate_Qb, BedLoad) and a steady-state
check that pins the momentum rows' friction and surface-slope terms.
Each check states an independent fact (an invariant, a hand-computed
number, a scaling law or a physical balance), chosen so that a swapped
index, a wrong neighbour or a wrong exponent fails it."""
import numpy as np
import pytest

import oracle_lib as ol
from hydist_c_amd import scenarios as S


def wet_box(M=16, N=12):
    """All-water rectangle: every compute cell wet, no islands."""
    sc = S.open_box(M, N, seed=3, n_islands=0)
    o = ol.Oracle(sc)
    o.init()
    kh = o.get("KHOUOT")
    assert (kh[2:N + 1, 2:M + 1] == 0).all()
    return sc, o


def test_smoothing_keeps_constant_and_linear_fields():
    # Normalize (support_funcs.cpp:761-782): weights sum to one, so a
    # constant field and, at cells whose four neighbours are wet, a linear
    # field are unchanged
    sc, o = wet_box()
    N, M = sc.N, sc.M
    ii, jj = np.meshgrid(np.arange(N + 3), np.arange(M + 3), indexing="ij")
    for f, kern in (("TU", "normalize_u"), ("TV", "normalize_v")):
        lin = 0.3 + 0.01 * ii - 0.02 * jj
        o.set(f, lin)
        o.kernel(kern)
        out = o.get(f)
        assert np.allclose(out[3:N, 3:M], lin[3:N, 3:M], rtol=0, atol=1e-14), f
        o.set(f, np.full(lin.shape, 0.7))
        o.kernel(kern)
        assert np.allclose(o.get(f)[2:N + 1, 2:M + 1], 0.7, atol=1e-15)


def test_smoothing_hand_case_with_a_dry_far_neighbour():
    # t_u: the two neighbours along i always count, those along j only when
    # wet.  v = 1, i-neighbours 2 and 4, j-neighbours 6 (wet) and 9 (dry):
    # 0.94 * 1 + 0.06 * (2 + 4 + 6) / 3 = 1.18 (heso = 0.94, loader.cpp)
    sc, o = wet_box()
    i, j = 6, 8
    tu = np.zeros(sc.shape)
    tu[i, j], tu[i - 1, j], tu[i + 1, j], tu[i, j - 1], tu[i, j + 1] = 1.0, 2.0, 4.0, 6.0, 9.0
    kh = o.get("KHOUOT")
    kh[i, j + 1] = 1
    o.set("KHOUOT", kh)
    o.set("TU", tu)
    o.kernel("normalize_u")
    assert o.get("TU")[i, j] == pytest.approx(1.18, abs=1e-14)


def test_vth_is_the_speed_at_the_cell_centre():
    # Find_VTH (sediment_transport.cpp:575-602): t_u averaged with the face
    # below (i-1), t_v with the face to the left (j-1)
    sc, o = wet_box()
    N, M = sc.N, sc.M
    ii, jj = np.meshgrid(np.arange(N + 3), np.arange(M + 3), indexing="ij")
    o.set("TU", 3.0 * ii)
    o.set("TV", 4.0 * jj)
    o.set("HMOI", np.full(sc.shape, 5.0))
    o.kernel("find_vth")
    vth = o.get("VTH")
    want = np.hypot(3.0 * (ii - 0.5), 4.0 * (jj - 0.5))
    assert np.allclose(vth[2:N + 1, 2:M + 1], want[2:N + 1, 2:M + 1], rtol=1e-14)
    o.set("HMOI", np.full(sc.shape, 0.01))  # below H_TINH = 0.02: dry
    o.kernel("find_vth")
    assert (o.get("VTH")[2:N + 1, 2:M + 1] == 0).all()


def test_diffusion_coefficient_scaling_and_clamp():
    # hesoK (:539-573): K = 5.93 sqrt(g) hbar |t| / Cz, Cz = 18 log(12 H / Ks),
    # clamped to [5, 100]; Kx averages h over j-1, j
    sc, o = wet_box()
    N, M = sc.N, sc.M
    c = ol.load_coeffs(sc.coeff_overrides)
    h = np.full(sc.shape, 4.0)
    o.set("H", h)
    o.set("HTDU", np.full(sc.shape, 4.0))
    o.set("HTDV", np.full(sc.shape, 4.0))
    cz = 18 * np.log(12 * 4.0 / c.Ks)
    for speed in (3.0, 6.0):
        o.set("TU", np.full(sc.shape, speed))
        o.set("TV", np.full(sc.shape, -speed))
        o.kernel("heso_k")
        k = 5.93 * np.sqrt(9.81) * 2.0 * speed / cz   # hbar = (4 + 4) / 4
        want = min(100.0, max(5.0, k))
        assert np.allclose(o.get("KX")[2:N, 2:M], want, rtol=1e-13)
        assert np.allclose(o.get("KY")[2:N, 2:M], want, rtol=1e-13)
    # depth varying along j: Kx averages h over (j-1, j), Ky over (i-1, i)
    jj = np.arange(M + 3)[None, :] * np.ones((N + 3, 1))
    hj = 4.0 + 0.25 * jj
    o.set("H", hj)
    o.set("TU", np.full(sc.shape, 2.0))
    o.set("TV", np.full(sc.shape, 2.0))
    o.kernel("heso_k")
    kx_want = 5.93 * np.sqrt(9.81) * (hj[:, 1:] + hj[:, :-1]) * 0.25 * 2.0 / cz
    ky_want = 5.93 * np.sqrt(9.81) * (2 * hj[:, 1:]) * 0.25 * 2.0 / cz
    assert np.allclose(o.get("KX")[2:N, 3:M], np.clip(kx_want, 5, 100)[2:N, 2:M - 1], rtol=1e-13)
    assert np.allclose(o.get("KY")[2:N, 3:M], np.clip(ky_want, 5, 100)[2:N, 2:M - 1], rtol=1e-13)
    o.set("TU", np.full(sc.shape, 1e-4))
    o.kernel("heso_k")
    assert np.allclose(o.get("KX")[2:N, 2:M], 5.0)
    o.set("TU", np.full(sc.shape, 1e3))
    o.kernel("heso_k")
    assert np.allclose(o.get("KX")[2:N, 2:M], 100.0)


def test_bed_load_flux_direction_threshold_and_exponent():
    # Calculate_Qb (intended form): Qb ~ Tx^2.1 along the velocity, Tx =
    # tau_b / tau_e - 1 above the threshold, 0 below
    sc, o = wet_box()
    N, M = sc.N, sc.M
    c = ol.load_coeffs(sc.coeff_overrides)
    o.set("HMOI", np.full(sc.shape, 0.1))  # below ghtoe: tau_e = Toe
    o.set("TU", np.full(sc.shape, 0.6))
    o.set("TV", np.full(sc.shape, 0.8))
    o.set("VTH", np.full(sc.shape, 1.0))
    res = {}
    for tx in (0.5, 1.0, 2.0, -0.5):
        fw = (tx + 1) * c.Toe / c.ro  # tau_b = ro * Fw * VTH^2
        o.set("FW", np.full(sc.shape, fw))
        o.kernel("calculate_qb")
        res[tx] = (o.get("QBX")[3:N, 3:M], o.get("QBY")[3:N, 3:M])
    qx, qy = res[1.0]
    assert np.allclose(qx / qy, 0.6 / 0.8, rtol=1e-14)
    assert np.allclose(res[2.0][0] / qx, 2.0 ** 2.1, rtol=1e-12)
    assert np.allclose(res[0.5][0] / qx, 0.5 ** 2.1, rtol=1e-12)
    assert (res[-0.5][0] == 0).all() and (res[-0.5][1] == 0).all()


def test_bed_change_is_the_divergence_of_the_bed_load():
    # BedLoad: dH += dT / (1 - Dorong) * (S + div Qb + diffusion); with FS = 0
    # (no source, no diffusion) and Qbx = a * i, Qby = b * j the interior
    # change is dT / (1 - Dorong) * (a / dX + b / dY)
    sc, o = wet_box(M=20, N=16)
    N, M = sc.N, sc.M
    c = ol.load_coeffs(sc.coeff_overrides)
    ii, jj = np.meshgrid(np.arange(N + 3), np.arange(M + 3), indexing="ij")
    a, b = 1e-3, -4e-4
    o.set("QBX", a * ii)
    o.set("QBY", b * jj)
    o.set("FS", np.zeros(sc.shape))
    o.set("DH", np.zeros(sc.shape))
    o.set("HMOI", np.full(sc.shape, 3.0))
    o.kernel("bed_load")
    want = c.dT / (1 - c.Dorong) * (a / c.dX + b / c.dY)
    dh = o.get("DH")
    assert np.allclose(dh[4:N - 2, 4:M - 2], want, rtol=1e-12), (dh[6, 6], want)


def test_steady_channel_balances_surface_slope_and_friction():
    # Steady flow in a flat straight channel: the momentum rows reduce to
    # g dz/dy = Ky1 |v| v / Htdv with Ky1 = g hbar^(-2 mu_mn) n^2
    # (support_funcs.cpp:64-67), and continuity gives v = Q / (width H).
    # Advection leaves a residual below one percent at these speeds.
    sc = S.straight_channel(M=80, N=8, depth=4.0, manning=0.03, Q=-100.0, hours=4)
    o = ol.Oracle(sc)
    o.init()
    o.step(5000)
    c = ol.load_coeffs(sc.coeff_overrides)
    z, v = o.get("Z"), o.get("V")
    ky1, htdv = o.get("KY1"), o.get("HTDV")
    i, js = 4, np.arange(20, 61)
    assert np.allclose(ky1[i, js], c.g * 4.0 ** (-2 * c.mu_mn) * 0.03 ** 2, rtol=1e-6)
    lhs = c.g * (z[i, js + 1] - z[i, js]) / c.dY
    rhs = -ky1[i, js] * np.abs(v[i, js]) * v[i, js] / htdv[i, js]
    assert np.all(np.abs(lhs / rhs - 1) < 0.02), (lhs / rhs).min()
    width = 5 * c.dX  # rows 3..N-1
    depth = 4.0 + z[i, js]
    assert np.allclose(-v[i, js], 100.0 / (width * depth), rtol=0.01)
