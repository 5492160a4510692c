"""Python binding of the gfx950 library (libhydist_amd.so) over its C ABI.

`Simulation` mirrors the reference's host driver: create (allocation and
upload, loader.cpp:224-375), init (main.cpp:257-299), step / run
(Hydraulic_Calculation, engine.cpp:41-296) and field access (save_result,
loader.cpp:397-414).  There is no CPU path: loading fails loudly when the
library is missing.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import F, MeshHolder, make_config, make_options

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhydist_amd.so")

_lib = None


class HydistError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HydistError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.hd_load_coeffs.argtypes = [C.POINTER(abi.Coeffs)]
        L.hd_finalize_coeffs.argtypes = [C.POINTER(abi.Coeffs)]
        L.hd_create.argtypes = [C.POINTER(abi.Mesh), C.POINTER(abi.Coeffs), C.POINTER(abi.Options),
                                C.POINTER(abi.Config), C.POINTER(vp)]
        for n in ("hd_init", "hd_sync", "hd_destroy"):
            getattr(L, n).argtypes = [vp]
        for n in ("hd_update_boundary", "hd_halfstep_uz", "hd_halfstep_vz", "hd_sediment_j", "hd_sediment_i",
                  "hd_bedload"):
            getattr(L, n).argtypes = [vp, C.c_double]
        L.hd_step.argtypes = [vp, dp]
        L.hd_run.argtypes = [vp, dp, C.c_double]
        for n in ("hd_copy_out", "hd_copy_in", "hd_copy_out_boundary", "hd_copy_in_boundary"):
            getattr(L, n).argtypes = [vp, C.c_int, dp, C.c_size_t]
        for n in ("hd_copy_out_i32", "hd_copy_in_i32"):
            getattr(L, n).argtypes = [vp, C.c_int, ip, C.c_size_t]
        L.hd_get_hmax.argtypes = [vp, dp]
        L.hd_set_hmax.argtypes = [vp, dp]
        L.hd_get_segments.argtypes = [vp, C.c_int, ip, C.c_size_t, ip, ip, C.c_size_t, C.POINTER(C.c_int64)]
        L.hd_mass_balance.argtypes = [vp, dp, dp]
        L.hd_last_error.argtypes = [vp]
        L.hd_last_error.restype = C.c_char_p
        L.hd_device_info.argtypes = [C.c_char_p, C.c_size_t]
        L.hd_batch_create.argtypes = [C.POINTER(abi.Mesh), C.c_int, C.POINTER(abi.Coeffs), C.POINTER(abi.Options),
                                      C.POINTER(abi.Config), C.POINTER(vp)]
        L.hd_batch_size.argtypes = [vp, C.POINTER(C.c_int)]
        for n in ("hd_batch_copy_out", "hd_batch_copy_in"):
            getattr(L, n).argtypes = [vp, C.c_int, C.c_int, dp, C.c_size_t]
        L.hd_set_graph.argtypes = [vp, C.c_int]
        L.hd_autotune.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.hd_set_solver_threads.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.hd_set_solver.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.hd_gtsv_batch.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.hd_snapshot_size.argtypes = [vp, C.POINTER(C.c_size_t)]
        L.hd_snapshot.argtypes = [vp, C.c_void_p, C.c_size_t, C.c_double]
        L.hd_restore.argtypes = [vp, C.c_void_p, C.c_size_t, dp]
        L.hd_save_checkpoint.argtypes = [vp, C.c_char_p, C.c_double]
        L.hd_load_checkpoint.argtypes = [vp, C.c_char_p, dp]
        L.hd_gtsv_long.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp]
        L.hd_gtsv_long_noalloc.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp]
        L.hd_set_watchdog.argtypes = [vp, C.c_int]
        L.hd_inject_fault.argtypes = [vp, C.c_int]
        _lib = L
    return _lib


def check(rc, handle=None):
    if rc != 0:
        msg = lib().hd_last_error(handle).decode(errors="replace")
        raise HydistError(f"hydist error {rc}: {msg}")


def device_info():
    buf = C.create_string_buffer(256)
    rc = lib().hd_device_info(buf, 256)
    return rc, buf.value.decode()


def load_coeffs(overrides=None):
    c = abi.Coeffs()
    check(lib().hd_load_coeffs(C.byref(c)))
    if overrides:
        for k, v in overrides.items():
            setattr(c, k, v)
        check(lib().hd_finalize_coeffs(C.byref(c)))
    return c


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Simulation:
    """One scenario resident on the GPU."""

    def __init__(self, scen, form=abi.HD_FORM_FP64_PLAIN, sediment=False, bed_change=False, quirks=0,
                 Tmax=3600.0, coeffs=None, device=0):
        self.scen = scen
        self.shape = scen.shape
        self.holder = MeshHolder(scen)
        self.coeffs = coeffs if coeffs is not None else load_coeffs(scen.coeff_overrides)
        self.ops = make_options(scen, Tmax)
        self.cfg = make_config(form=form, sediment=sediment, bed_change=bed_change, quirks=quirks, device=device)
        h = C.c_void_p()
        rc = lib().hd_create(C.byref(self.holder.mesh), C.byref(self.coeffs), C.byref(self.ops),
                             C.byref(self.cfg), C.byref(h))
        check(rc)
        self.h = h
        self.t = 0.0

    def close(self):
        if getattr(self, "h", None) is not None:
            lib().hd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self):
        check(lib().hd_init(self.h), self.h)

    def step(self, n=1):
        t = C.c_double(self.t)
        for _ in range(n):
            check(lib().hd_step(self.h, C.byref(t)), self.h)
        self.t = t.value

    def halfstep_uz(self, t):
        check(lib().hd_halfstep_uz(self.h, t), self.h)

    def stage(self, name, t):
        """Stage block: 'sediment_j', 'sediment_i' or 'bedload' (hd_sediment_j, ...)."""
        check(getattr(lib(), "hd_" + name)(self.h, t), self.h)

    def run(self, Tmax, interval=None, callback=None):
        """hd_run / hd_run_cb: steps until t >= int(Tmax); with a callback,
        callback(t) runs at every whole multiple of `interval` seconds and
        returning True stops the run."""
        t = C.c_double(self.t)
        if callback is None:
            rc = lib().hd_run(self.h, C.byref(t), Tmax)
            self.t = t.value  # the time reached, also when the run stopped early
            check(rc, self.h)
        else:
            CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_void_p)
            cb = CB(lambda h, tt, u: int(bool(callback(tt))))
            lib().hd_run_cb.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_int, CB, C.c_void_p]
            rc = lib().hd_run_cb(self.h, C.byref(t), Tmax, interval, cb, None)
            self.t = t.value
            if rc == abi.HD_E_CANCELLED:
                return False
            check(rc, self.h)
        self.t = t.value
        return True

    def halfstep_vz(self, t):
        check(lib().hd_halfstep_vz(self.h, t), self.h)

    def sync(self):
        check(lib().hd_sync(self.h), self.h)

    def get(self, name):
        if name == "KHOUOT":
            a = np.empty(self.shape, dtype=np.int32)
            check(lib().hd_copy_out_i32(self.h, F[name], a.ctypes.data_as(C.POINTER(C.c_int32)), a.size), self.h)
            return a
        a = np.empty(self.shape, dtype=np.float64)
        check(lib().hd_copy_out(self.h, F[name], _dp(a), a.size), self.h)
        return a

    def set(self, name, a):
        if name == "KHOUOT":
            a = np.ascontiguousarray(a, dtype=np.int32)
            check(lib().hd_copy_in_i32(self.h, F[name], a.ctypes.data_as(C.POINTER(C.c_int32)), a.size), self.h)
            return
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().hd_copy_in(self.h, F[name], _dp(a), a.size), self.h)

    def boundary_size(self, which):
        M, N = self.scen.M, self.scen.N
        return {"UBT": M + 3, "UBP": M + 3, "VBT": N + 3, "VBD": N + 3, "HI": 2 * (M + N + 6)}[which]

    def get_boundary(self, which):
        a = np.empty(self.boundary_size(which))
        check(lib().hd_copy_out_boundary(self.h, abi.BOUNDARY[which], _dp(a), a.size), self.h)
        return a

    def set_boundary(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().hd_copy_in_boundary(self.h, abi.BOUNDARY[which], _dp(a), a.size), self.h)

    def get_hmax(self):
        a = np.empty(4)
        check(lib().hd_get_hmax(self.h, _dp(a)), self.h)
        return a

    def set_hmax(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().hd_set_hmax(self.h, _dp(a)), self.h)

    def segments(self, direction):
        nlines = self.shape[0] if direction == 0 else self.shape[1]
        counts = np.zeros(nlines, dtype=np.int32)
        total = C.c_int64()
        ipc = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        check(lib().hd_get_segments(self.h, direction, ipc(counts), nlines, None, None, 0, C.byref(total)), self.h)
        n = total.value
        first = np.zeros(max(n, 1), dtype=np.int32)
        last = np.zeros(max(n, 1), dtype=np.int32)
        check(lib().hd_get_segments(self.h, direction, ipc(counts), nlines, ipc(first), ipc(last), n,
                                    C.byref(total)), self.h)
        return counts, first[:n], last[:n]

    def save_checkpoint(self, path):
        """Writes the full state and time (restart: continue bit for bit)."""
        check(lib().hd_save_checkpoint(self.h, str(path).encode(), self.t), self.h)

    def load_checkpoint(self, path):
        t = C.c_double()
        check(lib().hd_load_checkpoint(self.h, str(path).encode(), C.byref(t)), self.h)
        self.t = t.value

    def mass_balance(self):
        v, q = C.c_double(), C.c_double()
        check(lib().hd_mass_balance(self.h, C.byref(v), C.byref(q)), self.h)
        return v.value, q.value

    def set_watchdog(self, every):
        """NaN/Inf check of z, u, v every `every` steps (0 = off)."""
        check(lib().hd_set_watchdog(self.h, int(every)), self.h)

    def inject_fault(self, stage):
        """Arms a failure of the next UZ (1) or VZ (2) half-step; 0 clears."""
        check(lib().hd_inject_fault(self.h, int(stage)), self.h)


def autotune(sim):
    """Times the line-solver candidates of every sweep and keeps the fastest
    per sweep; returns the report ({candidate: ms, "chosen": [uz, vz, fs_j,
    fs_i]} with choices such as "part/4" or "cr/256")."""
    import json
    buf = C.create_string_buffer(2048)
    check(lib().hd_autotune(sim.h, buf, 2048), sim.h)
    return json.loads(buf.value.decode())


def set_solver_threads(sim, direction, hydraulic, threads):
    """Cyclic reduction with this workgroup size (0 = its default) for one sweep."""
    check(lib().hd_set_solver_threads(sim.h, direction, int(hydraulic), threads), sim.h)


SOLVER_DEFAULT, SOLVER_PARTITION, SOLVER_CR, SOLVER_SEGMENTS = 0, 1, 2, 3


def set_solver(sim, direction, hydraulic, kind, param=0):
    """Line solver of one sweep: SOLVER_PARTITION with K unknowns per thread
    (2, 4, 8; 0 = default), SOLVER_CR with a workgroup size, or
    SOLVER_DEFAULT; in the FP64 forms SOLVER_SEGMENTS (Thomas with one lane
    per system: in the plain form the same bits as the default, one lane
    per line; in the fast FP64 form one more choice at rounding level)."""
    check(lib().hd_set_solver(sim.h, direction, int(hydraulic), kind, param), sim.h)


def gtsv_batch(dl, d, du, b, x, m, nsys, ld, precision=0, stream=None):
    """hd_gtsv_batch on device pointers (ints); raises on error."""
    check(lib().hd_gtsv_batch(dl, d, du, b, x, m, nsys, ld, precision, stream), None)


def gtsv_long(dl, d, du, b, x, m, precision=0, stream=None):
    """hd_gtsv_long (one system of any length) on device pointers (ints);
    raises on error."""
    check(lib().hd_gtsv_long(dl, d, du, b, x, m, precision, stream), None)


def gtsv_long_noalloc(dl, d, du, b, x, m, precision=0, stream=None):
    """hd_gtsv_long_noalloc: as gtsv_long with the library's fixed work
    array (m <= GTSV_NOALLOC_MAX_M), no device allocation."""
    check(lib().hd_gtsv_long_noalloc(dl, d, du, b, x, m, precision, stream), None)


GTSV_NOALLOC_MAX_M = 262144


def _stage_api():
    L = lib()
    L.hd_set_profiling.argtypes = [C.c_void_p, C.c_int]
    L.hd_stage_stats.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.c_int]
    return L


def set_profiling(sim, on):
    check(_stage_api().hd_set_profiling(sim.h, int(on)), sim.h)


def stage_stats(sim, reset=True):
    """{stage: (launches, total_ms)} measured with HIP events on the
    simulation's stream."""
    import json
    buf = C.create_string_buffer(1 << 16)
    check(_stage_api().hd_stage_stats(sim.h, buf, len(buf), int(reset)), sim.h)
    return {k: (v[0], v[1]) for k, v in json.loads(buf.value.decode()).items()}


class Batch:
    """n scenarios of one mesh shape advanced together: every stage is one
    launch over all of them; optionally each step is one graph launch."""

    def __init__(self, scens, form=abi.HD_FORM_FP64_PLAIN, sediment=False, bed_change=False, quirks=0,
                 Tmax=3600.0, coeffs=None, device=0, graph=False):
        self.scens = list(scens)
        n = len(self.scens)
        self.shape = self.scens[0].shape
        self.holders = [MeshHolder(s) for s in self.scens]
        meshes = (abi.Mesh * n)(*[h.mesh for h in self.holders])
        self.coeffs = coeffs if coeffs is not None else load_coeffs(self.scens[0].coeff_overrides)
        self.ops = make_options(self.scens[0], Tmax)
        self.cfg = make_config(form=form, sediment=sediment, bed_change=bed_change, quirks=quirks, device=device)
        h = C.c_void_p()
        check(lib().hd_batch_create(meshes, n, C.byref(self.coeffs), C.byref(self.ops), C.byref(self.cfg),
                                    C.byref(h)))
        self.h = h
        self.n = n
        self.t = 0.0
        if graph:
            check(lib().hd_set_graph(self.h, 1), self.h)

    def close(self):
        if getattr(self, "h", None) is not None:
            lib().hd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self):
        check(lib().hd_init(self.h), self.h)

    def step(self, n=1):
        t = C.c_double(self.t)
        for _ in range(n):
            check(lib().hd_step(self.h, C.byref(t)), self.h)
        self.t = t.value

    def sync(self):
        check(lib().hd_sync(self.h), self.h)

    def get(self, scen, name):
        a = np.empty(self.shape, dtype=np.float64)
        check(lib().hd_batch_copy_out(self.h, scen, F[name], _dp(a), a.size), self.h)
        return a

    def set(self, scen, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().hd_batch_copy_in(self.h, scen, F[name], _dp(a), a.size), self.h)


class _ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("steps_per_quantum", C.c_int),
                ("use_graph", C.c_int), ("device", C.c_int)]


RUN_QUEUED, RUN_RUNNING, RUN_DONE, RUN_FAILED, RUN_CANCELLED = range(5)


def _server_lib():
    L = lib()
    if not getattr(L, "_server_bound", False):
        vp, i64 = C.c_void_p, C.c_int64
        L.hd_server_create.argtypes = [C.POINTER(_ServerConfig), C.POINTER(vp)]
        L.hd_server_submit.argtypes = [vp, C.POINTER(abi.Mesh), C.POINTER(abi.Coeffs), C.POINTER(abi.Options),
                                       C.POINTER(abi.Config), C.POINTER(i64)]
        L.hd_server_poll.argtypes = [vp, i64, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.hd_server_wait.argtypes = [vp, i64, C.c_double, C.POINTER(C.c_int)]
        L.hd_server_result.argtypes = [vp, i64, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.hd_server_error.argtypes = [vp, i64, C.c_char_p, C.c_size_t]
        for n in ("hd_server_cancel", "hd_server_release"):
            getattr(L, n).argtypes = [vp, i64]
        L.hd_server_stats.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.hd_server_destroy.argtypes = [vp]
        L.hd_footprint.argtypes = [C.POINTER(abi.Mesh), C.POINTER(abi.Config), C.c_int, C.POINTER(C.c_size_t)]
        L.hd_memory_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
        L._server_bound = True
    return L


def footprint(scen, nscen=1, dense_layout=0):
    L = _server_lib()
    holder = MeshHolder(scen)
    cfg = make_config()
    cfg.dense_layout = dense_layout
    out = C.c_size_t()
    check(L.hd_footprint(C.byref(holder.mesh), C.byref(cfg), nscen, C.byref(out)))
    return out.value


class Server:
    """The serving loop: submit what-if runs as they arrive; they are
    admitted by HBM footprint and advanced together on the GPU.  Queued
    runs of one mesh size, duration, form and options share one batched
    context (stats()["batched"] counts them)."""

    def __init__(self, hbm_budget=64 << 30, max_active=16, steps_per_quantum=16, use_graph=True, device=0):
        L = _server_lib()
        cfg = _ServerConfig(int(hbm_budget), max_active, steps_per_quantum, int(use_graph), device)
        h = C.c_void_p()
        check(L.hd_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.shapes = {}

    def submit(self, scen, Tmax, form=abi.HD_FORM_FP64_PLAIN, sediment=False, bed_change=False, coeffs=None):
        L = _server_lib()
        holder = MeshHolder(scen)
        c = coeffs if coeffs is not None else load_coeffs(scen.coeff_overrides)
        ops = make_options(scen, Tmax)
        cfg = make_config(form=form, sediment=sediment, bed_change=bed_change)
        rid = C.c_int64()
        check(L.hd_server_submit(self.h, C.byref(holder.mesh), C.byref(c), C.byref(ops), C.byref(cfg),
                                 C.byref(rid)))
        self.shapes[rid.value] = scen.shape
        return rid.value

    def poll(self, rid):
        st, t = C.c_int(), C.c_double()
        check(_server_lib().hd_server_poll(self.h, rid, C.byref(st), C.byref(t)))
        return st.value, t.value

    def wait(self, rid, timeout=-1.0):
        st = C.c_int()
        _server_lib().hd_server_wait(self.h, rid, timeout, C.byref(st))
        return st.value

    def result(self, rid, name):
        a = np.empty(self.shapes[rid], dtype=np.float64)
        rc = _server_lib().hd_server_result(self.h, rid, F[name], _dp(a), a.size)
        if rc != 0:
            raise HydistError(f"run {rid}: result unavailable ({rc}): {self.error(rid)}")
        return a

    def error(self, rid):
        buf = C.create_string_buffer(512)
        _server_lib().hd_server_error(self.h, rid, buf, 512)
        return buf.value.decode(errors="replace")

    def cancel(self, rid):
        check(_server_lib().hd_server_cancel(self.h, rid))

    def release(self, rid):
        check(_server_lib().hd_server_release(self.h, rid))

    def stats(self):
        import json
        buf = C.create_string_buffer(2048)
        check(_server_lib().hd_server_stats(self.h, buf, 2048))
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "h", None) is not None:
            _server_lib().hd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
