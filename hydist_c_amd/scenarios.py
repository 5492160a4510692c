"""Seeded scenario generators for the ADI timestep.

A scenario is the full input set of one run in the reference's padded
layout: (N+3) x (M+3) row-major arrays (rows i, columns j, unit stride j),
hourly boundary series per side and the sediment switches.

Depth convention (bandodosau): h is the bed depth below datum at cell
corners; a cell is wet when the mean of its four corners exceeds NANGDAY
(support_funcs.cpp:51).  Land is h = -3 (bed 3 m above datum).

The generators build corner depths so that every wet cell has four
positive corners and every partly-wet cell is dry.  Otherwise the
reference formulas take pow() of a negative face depth (Kx1/Ky1,
support_funcs.cpp:64-67; Q-boundary weights, :509-514) or divide by a zero
face depth, and produce NaN on the first step.
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np

LAND = -3.0


@dataclass
class Scenario:
    name: str
    M: int
    N: int
    h: np.ndarray
    hsnham: np.ndarray
    visc: np.ndarray
    Fw: np.ndarray
    total_time: int
    boundary_type: tuple
    bc: list
    CC: list = None
    sediment_start: float = 1e30
    bed_change_start: float = 1e30
    coeff_overrides: dict = field(default_factory=dict)
    u0: np.ndarray = None
    v0: np.ndarray = None
    z0: np.ndarray = None
    FS0: np.ndarray = None
    meta: dict = field(default_factory=dict)
    # Options members to override (e.g. the narrow-channel modes:
    # {"channel": True, "kenhhepd": True}, engine.cpp:24-30,77-80,190-193)
    options: dict = field(default_factory=dict)

    @property
    def shape(self):
        return (self.N + 3, self.M + 3)

    @property
    def cells(self):
        """Mesh cells counted by the cell-steps/s metric (all cells, wet or dry)."""
        return self.M * self.N


def _edge_pad(core):
    """standardize_inp.py's np.pad(..., ((1, 2), (1, 2)), 'edge')."""
    return np.pad(core, ((1, 2), (1, 2)), mode="edge")


def _wet_mask(h):
    """Onetime_init's mask for padded cells (i, j), i >= 1, j >= 1."""
    m = np.zeros(h.shape, dtype=bool)
    s = h[:-1, :-1] + h[:-1, 1:] + h[1:, :-1] + h[1:, 1:]
    m[1:, 1:] = s * 0.25 > 0.0
    return m


def build_bathymetry(M, N, in_channel_fn, depth_fn):
    """Corner depths on the padded grid from a cell-membership function and
    a positive corner depth profile.

    A corner is positive (depth_fn) when any cell sharing it is in the
    channel.  Each land corner is then set below minus the sum of the
    positive corners of every cell it touches, so a cell with a land corner
    always has a non-positive corner mean (dry), and a cell is wet exactly
    when its four corners are positive.  Depths on faces between a land
    corner and a positive corner are negative; such faces belong only to
    dry cells, whose coefficients the timestep never reads."""
    H, W = N + 3, M + 3
    ii, jj = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64), indexing="ij")
    # cell (i, j) centre in core coordinates (i - 1.5, j - 1.5); corner (a, b) at (a - 1, b - 1)
    cells = np.zeros((H, W), dtype=bool)
    cells[1:, 1:] = in_channel_fn(ii[1:, 1:] - 1.5, jj[1:, 1:] - 1.5)
    depth = np.maximum(depth_fn(ii - 1.0, jj - 1.0), 0.05)
    pos = np.zeros((H, W), dtype=bool)  # corner positivity; cell (i,j) owns corners (i-1..i, j-1..j)
    pos[:-1, :-1] |= cells[1:, 1:]
    pos[:-1, 1:] |= cells[1:, 1:]
    pos[1:, :-1] |= cells[1:, 1:]
    pos[1:, 1:] |= cells[1:, 1:]
    h = np.where(pos, depth, 0.0)
    # per cell: sum of its positive corners
    cs = np.zeros((H, W))
    cs[1:, 1:] = h[:-1, :-1] + h[:-1, 1:] + h[1:, :-1] + h[1:, 1:]
    # per corner: max over the (up to 4) cells sharing it
    need = np.zeros((H, W))
    need[:-1, :-1] = np.maximum(need[:-1, :-1], cs[1:, 1:])
    need[:-1, 1:] = np.maximum(need[:-1, 1:], cs[1:, 1:])
    need[1:, :-1] = np.maximum(need[1:, :-1], cs[1:, 1:])
    need[1:, 1:] = np.maximum(need[1:, 1:], cs[1:, 1:])
    land = np.minimum(LAND, -need - 0.01)
    h = np.where(pos, depth, land)
    h = _edge_pad(h[1:N + 1, 1:M + 1])
    wet = _wet_mask(h)
    allpos = np.zeros((H, W), dtype=bool)
    allpos[1:, 1:] = (h[:-1, :-1] > 0) & (h[:-1, 1:] > 0) & (h[1:, :-1] > 0) & (h[1:, 1:] > 0)
    assert (wet[2:N + 1, 2:M + 1] == allpos[2:N + 1, 2:M + 1]).all()
    return np.ascontiguousarray(h)


def _series(values, total_time):
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    if v.size == 1:
        v = np.full(total_time, float(v[0]))
    assert v.size >= total_time
    return np.ascontiguousarray(v[:total_time])


def meander_channel(M, N, seed=1, amplitude=20.0, wavelength=200.0, half_width=20.0,
                    depth_bank=2.0, depth_max=8.0, Q=-300.0, hours=1, manning=(0.025, 0.035),
                    visc=5.0, Fw=0.0, tide=None, manning_factor=1.0, name=None):
    """Config-1 family: sinusoidal channel along j with a parabolic section,
    land outside; up side (j = M) Q inflow, down side (j = 2) water level."""
    rng = np.random.default_rng(seed)
    yc0 = N / 2.0

    def centre(x):
        return yc0 + amplitude * np.sin(2 * np.pi * x / wavelength)

    def in_channel(y, x):
        return np.abs(y - centre(x)) <= half_width

    def depth(y, x):
        r = np.abs(y - centre(x)) / half_width
        return depth_bank + (depth_max - depth_bank) * np.maximum(0.0, 1.0 - r * r)

    h = build_bathymetry(M, N, in_channel, depth)
    shape = h.shape
    hsn = rng.uniform(manning[0], manning[1], size=shape) * manning_factor
    total_time = hours + 3
    bc_up = _series(Q, total_time)
    if tide is None:
        bc_down = _series(0.0, total_time)
    else:
        amp, period_h = tide
        t = np.arange(total_time, dtype=np.float64)
        bc_down = amp * np.sin(2 * np.pi * t / period_h)
    zeros = _series(0.0, total_time)
    return Scenario(
        name=name or f"meander_{M}x{N}", M=M, N=N, h=h, hsnham=hsn,
        visc=np.full(shape, visc), Fw=np.full(shape, Fw), total_time=total_time,
        boundary_type=(2, 0, 0, 0), bc=[bc_up, bc_down, zeros, zeros], CC=None,
        meta={"seed": seed, "Q": Q},
    )


def config1(seed=1):
    """256 x 128, dx = 10 m: amplitude 20, wavelength 200, width 40 cells,
    depth 2-8 m, Manning U[0.025, 0.035], A = 5, Fw = 0; Q = -300 m^3/s in at
    the up side, z = 0 at the down side; 1 simulated hour."""
    return meander_channel(256, 128, seed=seed, name="config1_256x128")


def meander_islands(M, N, seed=2, n_islands=30, amplitude=None, wavelength=None,
                    half_width=None, depth_bank=1.0, depth_max=15.0, Q=-4000.0, hours=1,
                    fs_range=(0.05, 0.2), Fw=0.0025, sediment_start=300.0, name=None):
    """Config-2 family: wide meandering channel with elliptic islands
    (radius 5-40 cells), depth 1-15 m tapering to 1 m at banks and island
    shores; up side Q inflow with FS concentration U[0.05, 0.2] kg/m^3,
    down side water level."""
    rng = np.random.default_rng(seed)
    amplitude = amplitude if amplitude is not None else 0.25 * N
    wavelength = wavelength if wavelength is not None else 0.5 * M
    half_width = half_width if half_width is not None else 0.24 * N
    yc0 = N / 2.0

    def centre(x):
        return yc0 + amplitude * np.sin(2 * np.pi * x / wavelength)

    islands = []
    while len(islands) < n_islands:
        a, b = rng.uniform(5, 40, size=2)
        cx = rng.uniform(0.05 * M, 0.95 * M)
        cy = centre(cx) + rng.uniform(-1, 1) * max(half_width - max(a, b) - 10, 1.0)
        islands.append((cy, cx, b, a))  # (row centre, col centre, row radius, col radius)

    # Only values below `cut` change the mesh: rho > 0 is all in_channel
    # asks, and a shore depth of depth_bank + 0.3 * cut already exceeds
    # depth_max.  Each island is therefore evaluated on the window of the
    # (meshgrid) coordinates where its rho can be below `cut`; elsewhere rho
    # stays +inf, which gives the same mesh, bit for bit, as evaluating
    # every island everywhere.
    cut = (depth_max - depth_bank) / 0.3 + 1.0

    def island_rho(y, x):
        y, x = np.broadcast_arrays(y, x)
        rho = np.full(y.shape, np.inf)
        ys, xs = y[:, 0], x[0, :]  # meshgrid coordinates: rows along y, columns along x
        for cy, cx, ry, rx in islands:
            R = 1.0 + cut / min(ry, rx)
            r0, r1 = np.searchsorted(ys, cy - R * ry, "left"), np.searchsorted(ys, cy + R * ry, "right")
            c0, c1 = np.searchsorted(xs, cx - R * rx, "left"), np.searchsorted(xs, cx + R * rx, "right")
            if r0 >= r1 or c0 >= c1:
                continue
            yy, xx = y[r0:r1, c0:c1], x[r0:r1, c0:c1]
            r = np.sqrt(((yy - cy) / ry) ** 2 + ((xx - cx) / rx) ** 2)
            rho[r0:r1, c0:c1] = np.minimum(rho[r0:r1, c0:c1], (r - 1.0) * min(ry, rx))
        return rho  # distance-like: < 0 inside an island

    def in_channel(y, x):
        return (np.abs(y - centre(x)) <= half_width) & (island_rho(y, x) > 0)

    def depth(y, x):
        r = np.abs(y - centre(x)) / half_width
        d = depth_bank + (depth_max - depth_bank) * np.maximum(0.0, 1.0 - r * r)
        shore = depth_bank + 0.3 * np.maximum(island_rho(y, x), 0.0)
        return np.minimum(d, shore)

    h = build_bathymetry(M, N, in_channel, depth)
    shape = h.shape
    hsn = rng.uniform(0.025, 0.035, size=shape)
    total_time = hours + 3
    zeros = _series(0.0, total_time)
    cc_up = rng.uniform(fs_range[0], fs_range[1], size=total_time)
    return Scenario(
        name=name or f"islands_{M}x{N}", M=M, N=N, h=h, hsnham=hsn,
        visc=np.full(shape, 5.0), Fw=np.full(shape, Fw), total_time=total_time,
        boundary_type=(2, 0, 0, 0), bc=[_series(Q, total_time), zeros, zeros, zeros],
        CC=[cc_up, zeros, zeros, zeros], sediment_start=sediment_start,
        bed_change_start=sediment_start,
        coeff_overrides={"dm": 2e-4, "Toe": 0.1, "Tod": 0.06},
        meta={"seed": seed, "islands": len(islands), "Q": Q},
    )


def config2(seed=2):
    """2048 x 1024 meander + 30 islands, sediment from t = 300 s, 1 hour."""
    return meander_islands(2048, 1024, seed=seed, name="config2_2048x1024")


def open_box(M, N, seed=3, n_islands=3, depth=(3.0, 6.0), Q_up=-150.0, Q_left=-80.0,
             z_down=0.05, z_right=0.0, types=(2, 0, 2, 0), hours=1, fs=True, name=None):
    """Small fully-open basin with islands: every side carries a boundary, so
    both sweeps see Q starts/ends, water-level starts/ends and wall ends."""
    rng = np.random.default_rng(seed)
    islands = []
    for _ in range(n_islands):
        cy = rng.uniform(0.25 * N, 0.75 * N)
        cx = rng.uniform(0.25 * M, 0.75 * M)
        ry, rx = rng.uniform(1.5, max(2.0, 0.12 * N)), rng.uniform(1.5, max(2.0, 0.12 * M))
        islands.append((cy, cx, ry, rx))

    def island_rho(y, x):
        rho = np.full(np.broadcast(y, x).shape, np.inf)
        for cy, cx, ry, rx in islands:
            r = np.sqrt(((y - cy) / ry) ** 2 + ((x - cx) / rx) ** 2)
            rho = np.minimum(rho, (r - 1.0) * min(ry, rx))
        return rho

    base = rng.uniform(depth[0], depth[1])

    def in_channel(y, x):
        return island_rho(y, x) > 0

    def dep(y, x):
        d = base + (depth[1] - depth[0]) * 0.5 * (1 + np.sin(0.3 * x) * np.cos(0.2 * y))
        shore = 1.0 + 0.5 * np.maximum(island_rho(y, x), 0.0)
        return np.minimum(d, shore)

    h = build_bathymetry(M, N, in_channel, dep)
    shape = h.shape
    total_time = hours + 3
    t = np.arange(total_time, dtype=np.float64)
    # per side: a discharge series when the side is Q-type (2), else a
    # water-level series
    q_mag = [Q_up, -Q_up, Q_left, -Q_left]
    z_mag = [0.03, z_down, 0.02, z_right + 0.01]
    bcs = []
    for side in range(4):
        if types[side] == 2:
            bcs.append(np.full(total_time, q_mag[side]) * (1 + 0.1 * np.sin(t + side)))
        else:
            bcs.append(z_mag[side] * np.cos(0.7 * t + side))
    bcs = [np.ascontiguousarray(b, dtype=np.float64) for b in bcs]
    cc = [rng.uniform(0.05, 0.2, size=total_time) for _ in range(4)] if fs else None
    return Scenario(
        name=name or f"openbox_{M}x{N}", M=M, N=N, h=h,
        hsnham=rng.uniform(0.02, 0.04, size=shape), visc=np.full(shape, 5.0),
        Fw=np.full(shape, 0.003), total_time=total_time, boundary_type=tuple(types),
        bc=bcs, CC=cc, sediment_start=0.0, bed_change_start=0.0,
        coeff_overrides={"dm": 2e-4, "Toe": 0.1, "Tod": 0.06},
        meta={"seed": seed},
    )


def closed_basin(M, N, seed=4, depth=4.0, z0_amp=0.0, name=None):
    """Basin with land on every side (no boundary segments): with z0 = 0 it
    is a lake at rest; with a bump in z0 it sloshes and conserves volume."""
    rng = np.random.default_rng(seed)

    def in_channel(y, x):
        return (y > 2) & (y < N - 4) & (x > 2) & (x < M - 4)

    def dep(y, x):
        return np.full(np.broadcast(y, x).shape, depth)

    h = build_bathymetry(M, N, in_channel, dep)
    shape = h.shape
    total_time = 4
    zeros = _series(0.0, total_time)
    z0 = None
    if z0_amp:
        ii, jj = np.meshgrid(np.arange(N + 3), np.arange(M + 3), indexing="ij")
        z0 = z0_amp * np.exp(-(((ii - N / 2) / (N / 6)) ** 2 + ((jj - M / 2) / (M / 6)) ** 2))
        z0 = np.ascontiguousarray(z0)
    return Scenario(
        name=name or f"basin_{M}x{N}", M=M, N=N, h=h,
        hsnham=rng.uniform(0.025, 0.035, size=shape), visc=np.full(shape, 5.0),
        Fw=np.zeros(shape), total_time=total_time, boundary_type=(0, 0, 0, 0),
        bc=[zeros, zeros, zeros, zeros], CC=None, z0=z0, meta={"seed": seed},
    )


def braided(M, N, seed=5, n_channels=12, width=(10, 120), depth=(1.0, 10.0), Q=-2000.0,
            hours=1, name=None):
    """Config-3 family: braided river of random-walk channels running along
    j with widths drawn from `width` (cells)."""
    rng = np.random.default_rng(seed)
    paths = []
    xs = np.arange(M + 3, dtype=np.float64)
    for _ in range(n_channels):
        y = rng.uniform(0.15 * N, 0.85 * N)
        steps = rng.normal(0, 1.5, size=M + 3)
        yc = np.clip(y + np.cumsum(steps), 0.1 * N, 0.9 * N)
        # smooth the walk
        k = np.ones(31) / 31.0
        yc = np.convolve(np.pad(yc, 15, mode="edge"), k, mode="valid")
        w = rng.uniform(width[0], width[1]) / 2.0
        paths.append((yc, w))

    def dist(y, x):
        xi = np.clip(np.rint(x).astype(np.int64), 0, M + 2)
        d = np.full(np.broadcast(y, x).shape, np.inf)
        for yc, w in paths:
            d = np.minimum(d, np.abs(y - yc[xi]) / w)
        return d  # <= 1 inside some channel

    def in_channel(y, x):
        return dist(y, x) <= 1.0

    def dep(y, x):
        r = dist(y, x)
        return depth[0] + (depth[1] - depth[0]) * np.maximum(0.0, 1.0 - r * r)

    h = build_bathymetry(M, N, in_channel, dep)
    shape = h.shape
    total_time = hours + 3
    zeros = _series(0.0, total_time)
    return Scenario(
        name=name or f"braided_{M}x{N}", M=M, N=N, h=h,
        hsnham=rng.uniform(0.025, 0.035, size=shape), visc=np.full(shape, 5.0),
        Fw=np.full(shape, 0.0025), total_time=total_time, boundary_type=(2, 0, 0, 0),
        bc=[_series(Q, total_time), zeros, zeros, zeros], CC=None, meta={"seed": seed},
    )


def config3(seed=5, sediment=False):
    """4096 x 4096 braided river (12 random-walk channels), dt = 1 s.  With
    sediment: suspended sediment from t = 0 with config 2's sediment model
    (d50 = 0.2 mm, Toe = 0.1, Tod = 0.06 N/m^2) and an up-side FS
    concentration U[0.05, 0.2] kg/m^3."""
    s = braided(4096, 4096, seed=seed, name="config3_4096x4096")
    s.coeff_overrides = dict(s.coeff_overrides, dT=1.0)
    if sediment:
        rng = np.random.default_rng(seed + 500)
        zeros = _series(0.0, s.total_time)
        s.CC = [rng.uniform(0.05, 0.2, size=s.total_time), zeros, zeros, zeros]
        s.coeff_overrides.update({"dm": 2e-4, "Toe": 0.1, "Tod": 0.06})
        s.sediment_start = 0.0
    return s


def ensemble_member(k, M=512, N=256):
    """Config-4 member k: the config-1 generator at 512 x 256 with seed k,
    inflow amplitude U[200, 1500] m^3/s, tide amplitude U[0.1, 0.8] m and a
    Manning factor U[0.8, 1.2]."""
    rng = np.random.default_rng(10_000 + k)
    Q = -rng.uniform(200, 1500)
    tide = rng.uniform(0.1, 0.8)
    mf = rng.uniform(0.8, 1.2)
    s = meander_channel(M, N, seed=k, amplitude=40.0, wavelength=400.0, half_width=40.0,
                        Q=Q, tide=(tide, 12.42), manning_factor=mf, name=f"member{k}_{M}x{N}")
    s.meta.update({"tide": tide, "manning_factor": mf})
    return s


def segment_lengths(h, M, N):
    """Wet-segment lengths of a corner-depth field under the reference's
    segmentation rules (Find_Calculation_limits_Horizontal/Vertical,
    support_funcs.cpp:238-343: a line is scanned from 2 to its end, a run
    starting at end - 1 or later is dropped, a run wet through the end
    includes it), for generator reports: {"rows": lengths along j of the
    lines i = 2..N, "cols": lengths along i of the lines j = 2..M}."""
    wet = _wet_mask(h)
    w = wet[2:N + 1, 2:M + 1]  # positions 2..end of every line
    out = {}
    for key, lines, end in (("rows", w, M), ("cols", w.T, N)):
        pad = np.zeros((lines.shape[0], lines.shape[1] + 2), dtype=np.int8)
        pad[:, 1:-1] = lines
        d = np.diff(pad, axis=1)
        s0 = np.nonzero(d == 1)[1]   # run starts, as position - 2
        e0 = np.nonzero(d == -1)[1]  # one past run ends, as position - 1
        start, last = s0 + 2, e0 + 1
        keep = start <= end - 2
        out[key] = (last - start + 1)[keep]
    return out


def segment_histogram(h, M, N, edges=(1, 3, 10, 100, 1000, 10000)):
    """Segment counts per length bin [edges[k], edges[k+1]) for rows and
    columns (segment_lengths)."""
    lens = segment_lengths(h, M, N)
    return {k: np.histogram(v, bins=edges)[0] for k, v in lens.items()}


def tidal_flats(M, N, seed=6, shelf=(0.03, 0.16), channel_depth=3.0, drop=2.0, rise_after=True,
                name=None):
    """Channel along j flanked by a shallow shelf (depth `shelf`); the down
    side (j = 2) water level falls by `drop` m in the first hour (and rises
    back in the second), so shelf cells dry out and later re-wet
    (Reset_states_vertical, support_funcs.cpp:142-188)."""
    rng = np.random.default_rng(seed)
    yc = N / 2.0

    def in_channel(y, x):
        return (y > 2) & (y < N - 3)

    def dep(y, x):
        r = np.abs(y - yc) / (0.5 * N)
        shelf_d = shelf[0] + (shelf[1] - shelf[0]) * rng.uniform(size=np.shape(y))
        return np.where(r < 0.35, channel_depth * (1 - r / 0.35) + shelf[1] * (r / 0.35), shelf_d)

    h = build_bathymetry(M, N, in_channel, dep)
    shape = h.shape
    total_time = 5
    z_down = np.array([0.0, -drop, 0.0 if rise_after else -drop, 0.0, 0.0])
    zeros = _series(0.0, total_time)
    return Scenario(
        name=name or f"flats_{M}x{N}", M=M, N=N, h=h, hsnham=rng.uniform(0.025, 0.035, size=shape),
        visc=np.full(shape, 5.0), Fw=np.full(shape, 0.003), total_time=total_time,
        boundary_type=(0, 0, 0, 0), bc=[zeros, np.ascontiguousarray(z_down), zeros, zeros], CC=None,
        meta={"seed": seed},
    )


# A request's tolerance on the water level (m) is log-uniform in
# [TOL_RANGE]; FP32 + refine serves requests whose tolerance is at least
# FP32_MIN_TOL (its error after one simulated hour on the 2048 x 1024
# mesh is a few 1e-12 m), FP64 the stricter half.
TOL_RANGE = (1e-12, 1e-4)
FP32_MIN_TOL = 1e-8


def form_for_tolerance(tol, fp64_form=0):
    return 2 if tol >= FP32_MIN_TOL else fp64_form


def serving_request(k, side=(128, 2048), duration=(600.0, 21600.0), p_sediment=0.5, fp64_form=0):
    """What-if request k of the serving workload: a meander + islands mesh
    with a log-uniform side in `side` (the other side 0.5-1.0 of it), a
    log-uniform simulated duration in `duration` (s), sediment on with
    probability p_sediment, and a log-uniform tolerance in TOL_RANGE that
    picks FP32 + refine (tolerance >= FP32_MIN_TOL, half the requests) or
    FP64 in form `fp64_form` (0 plain, 1 fast).
    Returns (scenario, Tmax, form, sediment)."""
    rng = np.random.default_rng(50_000 + k)
    M = int(round(np.exp(rng.uniform(np.log(side[0]), np.log(side[1])))))
    N = max(32, int(round(M * rng.uniform(0.5, 1.0))))
    T = float(np.exp(rng.uniform(np.log(duration[0]), np.log(duration[1]))))
    T = 2.0 * np.ceil(T / 2.0)
    sed = bool(rng.random() < p_sediment)
    tol = float(np.exp(rng.uniform(np.log(TOL_RANGE[0]), np.log(TOL_RANGE[1]))))
    form = form_for_tolerance(tol, fp64_form)
    n_isl = max(1, int(30 * M * N / (2048 * 1024)))
    s = meander_islands(M, N, seed=k, n_islands=n_isl, Q=-4000.0 * N / 1024, hours=int(np.ceil(T / 3600.0)),
                        sediment_start=0.0, name=f"request{k}_{M}x{N}")
    s.meta["tolerance"] = tol
    return s, T, form, sed


def straight_channel(M=160, N=8, depth=4.0, manning=0.03, Q=-10.0, z_down=0.0, hours=4, visc=5.0,
                     name=None):
    """Flat-bed straight channel along j, wet on rows 3..N-1 (rows 2 and N
    land, so the side lines carry no boundary): discharge Q enters on the up
    side (j = M), the down side (j = 2) holds z_down.  The steady state has
    a water surface whose slope balances bed friction."""
    def in_channel(y, x):
        return (y > 1.0) & (y < N - 2.0)

    def dep(y, x):
        return np.full(np.broadcast(y, x).shape, depth)

    h = build_bathymetry(M, N, in_channel, dep)
    shape = h.shape
    total_time = hours + 3
    zeros = _series(0.0, total_time)
    return Scenario(
        name=name or f"straight_{M}x{N}", M=M, N=N, h=h, hsnham=np.full(shape, manning),
        visc=np.full(shape, visc), Fw=np.zeros(shape), total_time=total_time, boundary_type=(2, 0, 0, 0),
        bc=[_series(Q, total_time), _series(z_down, total_time), zeros, zeros], CC=None,
        meta={"Q": Q, "manning": manning, "depth": depth},
    )


CONFIG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")


def load_config(name):
    """The benchmark configuration `name` (c1 ... c5, or the path of a JSON
    file of the same form): the generator and its arguments, Load_coeffs
    overrides (Constant_Coeffs member names), Options values (Tmax,
    sediment_start, bed_change_start), the forms it is run in and its
    sediment switches."""
    path = name if os.path.exists(name) else os.path.join(CONFIG_DIR, f"{name.lower()}.json")
    with open(path) as f:
        return json.load(f)


def config_scenario(cfg, k=0):
    """Scenario k of a configuration (a load_config result): the named
    generator called with the file's arguments (member k of an ensemble;
    request k of the serving workload, which returns (scenario, Tmax, form,
    sediment)), with the file's coefficient overrides and option values
    applied."""
    gen = globals()[cfg["generator"]]
    args = {a: tuple(v) if isinstance(v, list) else v for a, v in cfg.get("args", {}).items()}
    out = gen(k, **args) if cfg["generator"] in ("ensemble_member", "serving_request") else gen(**args)
    sc = out[0] if isinstance(out, tuple) else out
    sc.coeff_overrides = dict(sc.coeff_overrides, **cfg.get("coeffs", {}))
    opts = cfg.get("options", {})
    if "sediment_start" in opts:
        sc.sediment_start = float(opts["sediment_start"])
    if "bed_change_start" in opts:
        sc.bed_change_start = float(opts["bed_change_start"])
    return out

