"""Benchmark: ADI cell-steps/s of the Hydist_C timestep on one MI355X.

Workload: the 2048 x 1024 meander + 30 islands mesh with suspended
sediment and FP64 bed change on (both from t = 300 s), dt = 2 s.  A step is
one full dT (both ADI half-steps, engine.cpp:66,176); cell-steps/s counts
all M*N mesh cells.  State stays resident in HBM; the timed region starts
after the sediment switch-on so every timed step runs the full path.

Prints one JSON line (see the keys below).  `roofline` is measured for the
dominant kernel from HIP events on the library's stream; `cpu_baseline`
times the CPU oracle on this host's cores on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ADI cell-steps/s at 2048x1024, FP64 vs FP32+refine; p50/p99 scenario latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per mesh cell for one full step (both half-steps),
# FP64 state: hydraulics 204 B, + suspended sediment 48 B, + bed change 32 B.
STEP_BYTES = {"hydraulics": 204.0, "sediment": 252.0, "bed_change": 284.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def line_units(sim, sc):
    """Counts over the current wet-segment tables that the line kernels'
    algorithmic bytes are priced on.  For direction d (0: row lines, the
    VZ sweep and FS along i; 1: column lines, UZ and FS along j), over the
    segments of L cells:
      wet_d   sum L        positions whose state a line kernel reads
      il_d    sum 2L - 1   unknowns of the interleaved (z, velocity) systems
      fs_d    sum L - 2    FS unknowns (segments of 3 or more cells)
      fsc_d   sum L        cells of those FS segments (extraction writes)
    and pos_d, the line positions of a sweep over the whole mesh
    (identity rows included), for the all-positions figures."""
    u = {}
    for d in (0, 1):
        counts, first, last = sim.segments(d)
        L = (last - first + 1).astype("int64")
        long_ = L[L >= 3]
        u[f"wet_{d}"] = int(L.sum())
        u[f"il_{d}"] = int((2 * L - 1).sum())
        u[f"fs_{d}"] = int((long_ - 2).sum())
        u[f"fsc_{d}"] = int(long_.sum())
        u[f"pos_{d}"] = (sc.M - 1) * (sc.N - 1)
    return u


# Algorithmic bytes per unit of the line kernels: (unit, bytes).  FP64
# values: the hydraulic assembly reads 12 state/static fields and writes 2
# rows x 4 coefficients; a solve reads 4 coefficients and writes 1 value
# per unknown; extraction reads 2 solution values and writes t_z and the
# velocity; FS assembly reads 12 fields and writes 4 coefficients; FS
# extraction reads the solution and FS and writes tFS.  The one-kernel
# row-line sweeps read the assembly's fields and write the extraction's
# outputs: VZ 12 in + 2 out per wet position; FS along i 12 in per FS
# unknown + tFS out per cell of the segment.
LINE_PRICES = {
    "hyd_assemble": ("wet", 160.0),
    "hyd_solve": ("il", 40.0),
    "hyd_extract": ("wet", 32.0),
    "vz_fused": ("wet", 112.0),
    "fs_assemble": ("fs", 128.0),
    "fs_solve": ("fs", 40.0),
    "fs_extract": ("fs", 24.0),
}
# the same per position of every line (identity rows included): the
# all-positions figures of the row and column sweeps
ALL_POS_PRICES = {"hyd_assemble": 160.0, "hyd_solve": 80.0, "hyd_extract": 32.0, "vz_fused": 112.0,
                  "fs_assemble": 128.0, "fs_solve": 40.0, "fs_extract": 24.0, "fs_i_fused": 104.0}


def line_kind(name):
    """(kind, direction) of a line-kernel stage name, or None."""
    if not name.startswith(("k_cr", "k_part", "k_thomas", "k_hyd_assemble", "k_hyd_extract", "k_fs_assemble",
                            "k_fs_extract")):
        return None
    d = 1 if any(t in name for t in ("uz", "fs_j", "<j>")) else 0
    hyd = "fs" not in name
    if name.startswith("k_part<") and name.endswith(",vz>"):
        return "vz_fused", 0
    if name.startswith("k_part<") and name.endswith(",fs_i>"):
        return "fs_i_fused", 0
    if name.startswith(("k_cr", "k_part", "k_thomas")):
        return ("hyd_solve" if hyd else "fs_solve"), d
    for k in ("hyd_assemble", "hyd_extract", "fs_assemble", "fs_extract"):
        if name.startswith("k_" + k):
            return k, d
    return None


def stage_bytes(name, units, sc, all_positions=False):
    """Algorithmic bytes of one launch of a stage (FP64 state).  Line
    kernels are priced on the wet segments (line_units); with
    all_positions, on every position of every line instead."""
    cells = (sc.M + 3) * (sc.N + 3)
    lk = line_kind(name)
    if lk is not None:
        kind, d = lk
        if all_positions:
            return ALL_POS_PRICES[kind] * units[f"pos_{d}"]
        if kind == "fs_i_fused":
            return 96.0 * units["fs_0"] + 8.0 * units["fsc_0"]
        unit, price = LINE_PRICES[kind]
        return price * units[f"{unit}_{d}"]
    per_cell = {
        "k_smooth": 8 + 4 + 8,        # t_vel + khouot in, tmp out
        "k_solve_v": 10 * 8 + 4 + 16,  # stencil fields in, t_v + t_u out
        "k_solve_u": 10 * 8 + 4 + 16,
        # with sediment the post kernels also form VTH, Kx, Ky (three
        # fields out; H_moi, and for k_post_b htaiz, in)
        "k_post_a+vth": 6 * 8 + 4 + 6 * 8 + 8 + 3 * 8,
        "k_post_b+vth": 5 * 8 + 5 * 8 + 2 * 8 + 3 * 8,
        "k_post_a": 6 * 8 + 4 + 6 * 8,
        "k_post_b": 5 * 8 + 5 * 8,
        "k_pre_reset": 3 * 8 + 4 + 16,
        "k_dry": 8 + 4,
        "k_vth_k": 9 * 8 + 3 * 8,
        "k_update_fs": 16,
        "k_qb": 7 * 8 + 4 + 16,
        "k_bedload": 12 * 8 + 4 + 16,
        "k_segment": 8,               # khouot in, once per direction (segment tables are O(lines))
    }
    if name.startswith("k_rewet"):
        # candidate mask words in (the marked rows' reads are data dependent)
        return 8.0 * ((sc.N >> 6) + 1) * sc.M
    for k, v in per_cell.items():
        if name.startswith(k):
            return float(v) * cells
    return None


SWEEP_ORDER = ("uz", "vz", "fs_j", "fs_i")  # order of the autotuner's "chosen" list


def kernel_symbol(stage, sc=None, chosen=None, layout=None):
    """The kernel (as rocprofv3 names it) behind a stage name; `chosen` is
    the autotuner's choice per sweep (e.g. ["part/4", "cr/256", ...]),
    `layout` the sweep buffers' layout (0 plain form, 1 the fast forms)."""
    for prefix, kind in (("k_part<", "part"), ("k_cr_reg<", "cr")):
        if stage.startswith(prefix) and sc is not None:
            prec, what = stage[len(prefix):-1].split(",")
            t = "double, false" if prec == "f64" else "float, true"
            il = "false" if what.startswith("fs") else "true"
            n = sc.N - 1 if what in ("uz", "fs_j") else sc.M - 1  # unknowns per line
            pick = chosen[SWEEP_ORDER.index(what)] if chosen else None
            param = int(pick.split("/")[1]) if pick and pick.startswith(kind + "/") else 0
            if kind == "part":
                k = param or (4 if n <= 2048 else 8)
                # row-line sweeps (VZ, FS along i) are one kernel: the solver
                # assembles its rows and extracts (EXTRACT, ASSEMBLE)
                ex = "true" if what in ("vz", "fs_i") else "false"
                return f"{stage[:len(prefix) - 1]}<{t}, {il}, {k}, {ex}, {ex}>"
            else:
                threads = param or (512 if n > 1024 else 256)
                kpt = -(-n // threads)
                k = 2 if kpt <= 2 else 4 if kpt <= 4 else 8
            return f"{stage[:len(prefix) - 1]}<{t}, {il}, {k}>"
    if stage.startswith("k_thomas_seg"):
        return "k_thomas_seg"
    if stage.startswith("k_thomas"):
        return "k_thomas"
    if stage.endswith("+vth"):  # k_post_a / k_post_b with the VTH, K pass folded in
        return stage[:-len("+vth")]
    # assembly / extraction: template <direction (1 = UZ / FS along j), layout
    # (0 = [rho][line] for the plain form, 1 = [line][rho])>
    # (column lines into [line][rho] are assembled by the LDS-tiled kernels)
    if layout == 1 and stage == "k_hyd_assemble<uz>":
        return "k_hyd_assemble_cols<1>"
    if layout == 1 and stage == "k_fs_assemble<j>":
        return "k_fs_assemble_cols<1>"
    if layout == 1 and stage == "k_hyd_extract<uz>":
        return "k_hyd_extract_cols"
    if layout == 1 and stage == "k_fs_extract<j>":
        return "k_fs_extract_cols"
    for prefix in ("k_hyd_assemble<", "k_hyd_extract<", "k_fs_assemble<", "k_fs_extract<"):
        if stage.startswith(prefix) and layout is not None:
            what = stage[len(prefix):-1]
            return f"{prefix}{1 if what in ('uz', 'j') else 0}, {layout}>"
    return stage


def time_form(sc, form, steps, warmup, profile_steps):
    from hydist_c_amd import engine
    sim = engine.Simulation(sc, form=form, sediment=True, bed_change=True, Tmax=3600.0)
    sim.init()
    # advance to the sediment switch-on, then the warm-up steps
    while sim.t < sc.sediment_start:
        sim.step()
    # line solver per sweep (fast forms: partition K or cyclic-reduction
    # workgroup size; plain form: Thomas per line or per system, the same
    # bits either way)
    tune = engine.autotune(sim)
    sim.step(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / steps
    # stage profile (separate pass: events add a little host work per launch)
    engine.set_profiling(sim, True)
    sim.step(profile_steps)
    sim.sync()
    stats = engine.stage_stats(sim)
    engine.set_profiling(sim, False)
    z = sim.get("Z")
    finite = bool((z == z).all())
    volume, inflow = sim.mass_balance()
    mass = {"t": sim.t, "volume_m3": volume, "inflow_m3": inflow}
    return sim, ms_per_step, stats, finite, tune, mass


def cpu_baseline(sc, seconds, threads=None):
    """The oracle (OpenMP over lines) on this host, built -O3 -march=native
    for this CPU (a timing copy; the checker is the -O2 -ffp-contract=off
    build), on the same mesh with every stage active (sediment and bed
    change from t = 0), for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from hydist_c_amd.abi import make_config
    import copy
    model = ol.use_native_build()
    s2 = copy.copy(sc)
    s2.sediment_start = 0.0
    s2.bed_change_start = 0.0
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    ol.set_threads(threads)
    o = ol.Oracle(s2, cfg=make_config(sediment=True, bed_change=True))
    o.init()
    o.step(1)  # first touch
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    n = max(1, min(400, int(seconds / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    return {"value": sc.M * sc.N * n / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(), "build": "g++ -O3 -march=native -fopenmp",
            "sample": f"{n} full steps (hydraulics + sediment + bed change) of the {sc.M}x{sc.N} mesh, "
                      f"{dt:.1f} s on {threads} threads"}


def serving_measure(seconds, rate, seed=0, max_active=32, side=(128, 512), duration=(60.0, 600.0)):
    """The serving loop under Poisson arrivals: `rate` requests/s for
    `seconds` s of what-if runs (meander + islands meshes with sides
    log-uniform in [128, 512], durations log-uniform in [60, 600] s
    simulated, sediment on for half, FP32 + refine for half), then waits
    for them to finish.  Reports p50/p99 completion latency and the
    aggregate cell-steps/s of completed runs."""
    import numpy as np
    from hydist_c_amd import engine, scenarios as S
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 2) + 10))
    arrivals = arrivals[arrivals < seconds]
    reqs = [S.serving_request(k, side=side, duration=duration, fp64_form=1) for k in range(len(arrivals))]
    srv = engine.Server(hbm_budget=64 << 30, max_active=max_active, steps_per_quantum=8, use_graph=False)
    t0 = time.perf_counter()
    ids = []
    for at, (sc, T, form, sed) in zip(arrivals, reqs):
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        ids.append(srv.submit(sc, T, form=form, sediment=sed, bed_change=sed))
    deadline = time.perf_counter() + 120.0
    for rid in ids:
        srv.wait(rid, timeout=max(0.0, deadline - time.perf_counter()))
    st = srv.stats()
    srv.close()
    st.update({"rate_per_s": rate, "window_s": seconds,
               "workload": f"meander+islands, side log-U[{side[0]},{side[1]}], duration log-U[{duration[0]:g},"
                           f"{duration[1]:g}] s simulated, sediment p=0.5, FP32+refine p=0.5 else FP64 (fast form)"})
    return st


def serving_full(seconds, rate, max_active=32, hbm_budget=200 << 30, queue_cap=4, seed=0):
    """The serving workload at its full size: Poisson arrivals at `rate`
    requests/s for `seconds` s of what-if runs (meander + islands meshes,
    side log-uniform in [128, 2048], simulated duration log-uniform in
    [600, 21600] s, sediment on for half, FP32 + refine for half and FP64
    (fast form) for the rest).  The offered load is far above what one GPU
    can serve, so most requests are still waiting when the window closes.
    A request's mesh is generated and submitted only while fewer than
    `queue_cap` submitted runs wait for admission; the rest wait, in
    arrival order, on the client side.  Latency runs from arrival (client
    clock) to the poll that sees the run done.  At the end of the window
    the unfinished runs are cancelled; reported: completions, p50/p99
    latency of the completed runs, cell-steps served in the window
    (completed runs plus the progress of the running ones) and the
    backlog."""
    import collections
    import numpy as np
    from hydist_c_amd import engine, scenarios as S
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 2) + 10))
    arrivals = arrivals[arrivals < seconds]
    srv = engine.Server(hbm_budget=hbm_budget, max_active=max_active, steps_per_quantum=8, use_graph=False)
    backlog = collections.deque()
    live = {}  # rid -> (arrival, cells, dT steps, Tmax)
    lat, served, failed, nxt = [], 0.0, 0, 0
    side_done = []
    t0 = time.perf_counter()
    last_log = 0.0
    while True:
        now = time.perf_counter() - t0
        if now >= seconds:
            break
        while nxt < len(arrivals) and arrivals[nxt] <= now:
            backlog.append(nxt)
            nxt += 1
        waiting = 0
        for rid, (at, cells, steps, T) in list(live.items()):
            st, _ = srv.poll(rid)
            if st == engine.RUN_DONE or st in (engine.RUN_FAILED, engine.RUN_CANCELLED):
                if st == engine.RUN_DONE:
                    lat.append(time.perf_counter() - t0 - at)
                    served += cells * steps
                    side_done.append(int(round(cells ** 0.5)))
                else:
                    failed += 1
                srv.release(rid)
                del live[rid]
            elif st == engine.RUN_QUEUED:
                waiting += 1
        while backlog and waiting < queue_cap and time.perf_counter() - t0 < seconds:
            k = backlog.popleft()
            sc, T, form, sed = S.serving_request(k, fp64_form=1)
            rid = srv.submit(sc, T, form=form, sediment=sed, bed_change=sed)
            live[rid] = (float(arrivals[k]), sc.M * sc.N, int(np.ceil(T / 2.0)), T)
            waiting += 1
        if now - last_log >= 15.0:
            last_log = now
            log(f"[serving-full] t={now:5.0f}s arrived {nxt} completed {len(lat)} live {len(live)} "
                f"client backlog {len(backlog)}")
        time.sleep(0.02)
    # progress of the runs still live, then cancel them
    running = 0
    for rid, (at, cells, steps, T) in live.items():
        st, t = srv.poll(rid)
        if st == engine.RUN_RUNNING:
            running += 1
            served += cells * min(steps, int(t / 2.0))
        srv.cancel(rid)
    srv.close()
    lat = np.array(lat)
    return {"window_s": seconds, "rate_per_s": rate, "arrived": int(nxt), "completed": int(len(lat)),
            "failed": failed, "running_at_end": running, "backlog_at_end": int(nxt - len(lat) - failed),
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "cell_steps_served": served, "cell_steps_per_s": served / seconds,
            "median_side_completed": float(np.median(side_done)) if side_done else None,
            "max_active": max_active, "hbm_budget": hbm_budget,
            "workload": "meander+islands, side log-U[128,2048], duration log-U[600,21600] s simulated, "
                        "sediment p=0.5, tolerance log-U[1e-12,1e-4] m: FP32+refine at >= 1e-8 (half), "
                        "else FP64 (fast form)"}


def ensemble_measure(n_scen, steps, forms=("plain", "fast", "fp32")):
    """The ensemble workload: n_scen independent 512 x 256 meandering
    channels (per-scenario inflow amplitude, tide and Manning factor) in
    one batched context, every stage one launch over all of them and each
    step one graph launch.  Reports cell-steps/s per form over `steps`
    steps after 5 warm-up steps."""
    from hydist_c_amd import abi, engine, scenarios as S
    ids = {"plain": abi.HD_FORM_FP64_PLAIN, "fast": abi.HD_FORM_FP64_FAST, "fp32": abi.HD_FORM_FP32_REFINE}
    scens = [S.ensemble_member(k) for k in range(n_scen)]
    out = {"scenarios": n_scen, "mesh": [scens[0].M, scens[0].N], "steps": steps, "graph": True}
    for f in forms:
        b = engine.Batch(scens, form=ids[f], graph=True)
        b.init()
        b.step(5)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        dt = time.perf_counter() - t0
        cells = n_scen * scens[0].M * scens[0].N
        out[f] = {"ms_per_step": dt * 1e3 / steps, "cell_steps_per_s": cells * steps / dt}
        b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--forms", default="plain,fast,fp32", help="comma list of plain,fast,fp32")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=0.0,
                    help="arrival window of the small-request serving measurement (0 = skip)")
    ap.add_argument("--serving-rate", type=float, default=20.0)
    ap.add_argument("--serving-full", type=float, default=60.0,
                    help="arrival window (s) of the serving workload at its full request distribution "
                         "(600 for the configured window; 0 = skip)")
    ap.add_argument("--ensemble", type=int, default=256,
                    help="scenarios of the 512x256 ensemble measurement (0 = skip)")
    ap.add_argument("--ensemble-steps", type=int, default=60)
    ap.add_argument("--sediment-from-start", action="store_true",
                    help="switch sediment and bed change on at t=0 (short profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this library has no multi-GPU path "
              "(the mesh would shard by row bands with halo exchange; not built).", file=sys.stderr)
        sys.exit(2)

    from hydist_c_amd import abi, scenarios as S
    forms = {"plain": abi.HD_FORM_FP64_PLAIN, "fast": abi.HD_FORM_FP64_FAST, "fp32": abi.HD_FORM_FP32_REFINE}
    names = [f for f in args.forms.split(",") if f]
    log("[bench] generating the 2048x1024 meander + islands mesh")
    sc = S.config2()
    if args.sediment_from_start:
        sc.sediment_start = sc.bed_change_start = 0.0
    results = {}
    best = None
    for name in names:
        log(f"[bench] form {name}: warm-up to t={sc.sediment_start:.0f}s + {args.warmup} steps, "
            f"then {args.steps} timed steps")
        sim, ms, stats, finite, tune, mass = time_form(sc, forms[name], args.steps, args.warmup,
                                                       args.profile_steps)
        cs = sc.M * sc.N / (ms * 1e-3)
        results[name] = {"cell_steps_per_s": cs, "ms_per_step": ms, "finite": finite, "stats": stats, "sim": sim,
                         "tune": tune, "mass": mass}
        log(f"[bench] form {name}: {ms:.3f} ms/step, {cs:.3e} cell-steps/s")
        if best is None or cs > results[best]["cell_steps_per_s"]:
            if best is not None:
                results[best]["sim"].close()
            best = name
        else:
            sim.close()
    r = results[best]
    sim = r["sim"]
    stats = r["stats"]
    units = line_units(sim, sc)
    chosen = r["tune"]["chosen"] if r["tune"] else None
    layout = 0 if best == "plain" else 1
    # dominant kernel symbol by total time (a symbol can serve several
    # stages, e.g. one cyclic-reduction kernel for both hydraulic sweeps)
    groups = {}
    for name, (launches, total_ms) in stats.items():
        if launches == 0 or total_ms <= 0:
            continue
        g = groups.setdefault(kernel_symbol(name, sc, chosen, layout), [0, 0.0, 0.0, 0.0, True, []])
        b = stage_bytes(name, units, sc)
        ball = stage_bytes(name, units, sc, all_positions=True)
        g[0] += launches
        g[1] += total_ms
        g[5].append(name)
        if b is None:
            g[4] = False
        else:
            g[2] += b * launches
            g[3] += ball * launches
    roofline = None
    if groups:
        dom_name, (launches, total_ms, total_bytes, total_all, known, stage_names) = max(
            groups.items(), key=lambda kv: kv[1][1])
        avg_ms = total_ms / launches
        if known:
            nbytes = total_bytes / launches
            achieved = nbytes / (avg_ms * 1e-3) / 1e9
            all_b = total_all / launches
            roofline = {"bound": "hbm", "kernel": dom_name, "stages": stage_names, "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "bytes_per_launch": nbytes, "avg_launch_ms": avg_ms, "launches": launches,
                        "priced_on": "wet segments (line kernels) / all cells (stencil kernels)",
                        "all_positions": {"bytes_per_launch": all_b,
                                          "achieved": all_b / (avg_ms * 1e-3) / 1e9,
                                          "frac": all_b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    # the tridiagonal solve kernel of the best form, at its own roofline:
    # the slowest pure hydraulic solve, priced at 40 B per interleaved
    # unknown of the wet segments; the one-kernel VZ sweep beside it
    tri = None
    solver = [(k, v) for k, v in stats.items()
              if k.startswith(("k_cr", "k_part", "k_thomas")) and "fs" not in k and v[0] > 0 and v[1] > 0]

    def solver_roof(k, launches, total_ms):
        b = stage_bytes(k, units, sc)
        ball = stage_bytes(k, units, sc, all_positions=True)
        t = total_ms / launches * 1e-3
        return {"kernel": k, "bytes_per_launch": b, "avg_launch_ms": t * 1e3, "achieved": b / t / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b / t / 1e9 / HBM_PEAK_GBS,
                "all_positions_frac": ball / t / 1e9 / HBM_PEAK_GBS}

    fused = [(k, v) for k, v in solver if k.startswith("k_part<") and k.endswith(",vz>")]
    pure = [(k, v) for k, v in solver if (k, v) not in fused]
    if pure:
        k, (launches, total_ms) = max(pure, key=lambda kv: kv[1][1])
        tri = solver_roof(k, launches, total_ms)
    if fused:
        k, (launches, total_ms) = fused[0]
        f = solver_roof(k, launches, total_ms)
        f["note"] = ("the VZ sweep in one kernel: rows assembled from the state, solved and t_z, t_v extracted "
                     "on chip; bytes = the 12 state/static fields in + t_z, t_v out per wet position")
        if tri is None:
            tri = f
        else:
            tri["fused_vz"] = f
    step_bytes = STEP_BYTES["bed_change"] * sc.M * sc.N
    step_roof = {"bytes_per_step": step_bytes, "achieved_GBs": step_bytes / (r["ms_per_step"] * 1e-3) / 1e9}
    step_roof["frac"] = step_roof["achieved_GBs"] / HBM_PEAK_GBS
    # domain volume of each form at the same time, against the plain form
    # (bit for bit the oracle's hydraulics)
    mass = {k: dict(v["mass"]) for k, v in results.items()}
    if "plain" in mass:
        v0 = mass["plain"]["volume_m3"]
        for k, v in mass.items():
            v["rel_diff_vs_plain"] = (v["volume_m3"] - v0) / v0 if v0 else None
    ensemble = None
    if args.ensemble > 0:
        log(f"[bench] ensemble: {args.ensemble} x 512x256 in one batched context, {args.ensemble_steps} steps per form")
        ensemble = ensemble_measure(args.ensemble, args.ensemble_steps)
        log("[bench] ensemble: " + ", ".join(f"{f} {ensemble[f]['ms_per_step']:.2f} ms/step "
                                             f"{ensemble[f]['cell_steps_per_s']:.3e} cell-steps/s"
                                             for f in ("plain", "fast", "fp32") if f in ensemble))
    serving = None
    if args.serving_seconds > 0:
        log(f"[bench] serving loop: {args.serving_rate:.0f} req/s for {args.serving_seconds:.0f} s")
        serving = serving_measure(args.serving_seconds, args.serving_rate)
        log(f"[bench] serving: p50 {serving['p50_s']:.3f} s, p99 {serving['p99_s']:.3f} s, "
            f"{serving['completed']}/{serving['submitted']} done, {serving['cell_steps_per_s']:.3e} cell-steps/s")
    full = None
    if args.serving_full > 0:
        log(f"[bench] full serving workload: {args.serving_rate:.0f} req/s for {args.serving_full:.0f} s")
        full = serving_full(args.serving_full, args.serving_rate)
        log(f"[bench] full serving: {full['completed']}/{full['arrived']} done, p50 {full['p50_s']} s, "
            f"p99 {full['p99_s']} s, {full['cell_steps_per_s']:.3e} cell-steps/s, backlog {full['backlog_at_end']}")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU oracle baseline")
        cpu = cpu_baseline(sc, args.cpu_seconds)
        log(f"[bench] cpu: {cpu['value']:.3e} cell-steps/s ({cpu['sample']}, {cpu['cpu_model']})")
        cpu1 = cpu_baseline(sc, args.cpu_seconds / 2, threads=1)
        log(f"[bench] cpu, 1 thread: {cpu1['value']:.3e} cell-steps/s ({cpu1['sample']})")
        cpu["single_thread"] = {"value": cpu1["value"], "sample": cpu1["sample"]}
    out = {
        "metric": METRIC,
        "value": r["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best in ("plain", "fast") else "f32+f64",
        "data": "synthetic (seeded generator: meander + 30 elliptic islands)",
        "config": {"workload": "2048x1024 meander+islands, sediment + FP64 bed change, dt=2s",
                   "M": sc.M, "N": sc.N, "form": best, "timed_from_t": sc.sediment_start},
        "forms": {k: {"cell_steps_per_s": v["cell_steps_per_s"], "ms_per_step": v["ms_per_step"],
                      "finite": v["finite"], "solver": v["tune"]["chosen"] if v["tune"] else None,
                      "profile_top": sorted(({"stage": a, "launches": b[0], "avg_ms": b[1] / max(b[0], 1)}
                                             for a, b in v["stats"].items()), key=lambda d: -d["avg_ms"])[:6]}
                  for k, v in results.items()},
        "roofline": roofline,
        "line_units": units,
        "step_roofline": step_roof,
        "tridiag_roofline": tri,
        "mass_balance": mass,
        "cpu_baseline": cpu,
        "serving": serving,
        "serving_full": full,
        "ensemble": ensemble,
        "profile_top": sorted(({"stage": k, "launches": v[0], "avg_ms": v[1] / max(v[0], 1)} for k, v in stats.items()),
                              key=lambda d: -d["avg_ms"] * 1)[:8],
        # every stage of the fastest form: launches and average ms per
        # launch in the profile pass (HIP events on the library's stream)
        "stages": {k: [v[0], round(v[1] / max(v[0], 1), 5)] for k, v in sorted(stats.items()) if v[0]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
