"""The C ABI's argument and state checks: every entry point answers a NULL
handle or buffer, a wrong size, an out-of-range field, side, direction or
scenario, and a call out of order with its status code (HD_E_INVALID,
HD_E_STATE), never a crash, and the context stays usable afterwards (the
reference aborts on any failed CUDA call; a drop-in library returns)."""
import ctypes as C

import numpy as np
import pytest

from hydist_c_amd import abi, engine, scenarios as S

pytestmark = pytest.mark.gpu

INVALID, STATE = -1, -5


def test_calls_out_of_order_and_bad_arguments():
    print("[case] C ABI argument checks", flush=True)
    L = engine.lib()
    engine._server_lib()
    sc = S.open_box(20, 14)
    g = engine.Simulation(sc)
    h = g.h
    n = (sc.M + 3) * (sc.N + 3)
    buf = np.zeros(n)
    ib = np.zeros(n, dtype=np.int32)
    dp = buf.ctypes.data_as(C.POINTER(C.c_double))
    ip = ib.ctypes.data_as(C.POINTER(C.c_int32))
    t = C.c_double(0.0)
    # before hd_init: the time loop, its stages, the boundary update and
    # the mass balance need the mask and weights init derives
    assert L.hd_step(h, C.byref(t)) == STATE
    assert L.hd_halfstep_uz(h, C.c_double(0.0)) == STATE
    assert L.hd_halfstep_vz(h, C.c_double(0.0)) == STATE
    assert L.hd_run(h, C.byref(t), C.c_double(4.0)) == STATE
    assert L.hd_update_boundary(h, C.c_double(0.0)) == STATE
    assert L.hd_mass_balance(h, C.byref(t), C.byref(t)) == STATE
    assert "hd_init" in engine.lib().hd_last_error(h).decode()
    # uploaded fields can be read before init
    assert L.hd_copy_out(h, abi.F["Z"], dp, n) == 0
    g.init()
    # NULL handles and outputs
    for fn in (L.hd_init, L.hd_sync):
        assert fn(None) == INVALID
    assert L.hd_destroy(None) == 0
    assert L.hd_step(None, C.byref(t)) == INVALID
    assert L.hd_step(h, None) == INVALID
    assert L.hd_run(h, None, C.c_double(4.0)) == INVALID
    assert L.hd_get_hmax(h, None) == INVALID
    assert L.hd_set_hmax(h, None) == INVALID
    assert L.hd_snapshot_size(h, None) == INVALID
    assert L.hd_memory_bytes(h, None) == INVALID
    assert L.hd_batch_size(h, None) == INVALID
    assert L.hd_stage_stats(h, None, 10, 0) == INVALID
    assert L.hd_device_info(None, 10) == INVALID
    assert L.hd_device_info(C.create_string_buffer(4), 0) == INVALID
    # field copies: buffer, size, field id and element type
    assert L.hd_copy_out(h, abi.F["Z"], None, n) == INVALID
    for m in (n - 1, n + 1, 0):
        assert L.hd_copy_out(h, abi.F["Z"], dp, m) == INVALID
        assert L.hd_copy_in(h, abi.F["Z"], dp, m) == INVALID
    for f in (-1, max(abi.F.values()) + 1):
        assert L.hd_copy_out(h, f, dp, n) == INVALID
    assert L.hd_copy_out(h, abi.F["KHOUOT"], dp, n) == INVALID
    assert L.hd_copy_out_i32(h, abi.F["Z"], ip, n) == INVALID
    assert L.hd_copy_in(h, abi.F["Z"], None, n) == INVALID
    assert L.hd_copy_in_i32(h, abi.F["KHOUOT"], None, n) == INVALID
    # boundary arrays: side and size
    for which, m in ((-1, 10), (5, 10), (0, 10 ** 6)):
        assert L.hd_copy_out_boundary(h, which, dp, m) == INVALID
    assert L.hd_copy_out_boundary(h, 0, None, 10) == INVALID
    # segments, watchdog, fault injection, batch scenario
    assert L.hd_get_segments(h, 2, None, 0, None, None, 0, None) == INVALID
    assert L.hd_set_watchdog(h, -1) == INVALID
    assert L.hd_inject_fault(h, 99) == INVALID
    for k in (-1, 1):
        assert L.hd_batch_copy_out(h, k, abi.F["Z"], dp, n) == INVALID
    # snapshots: short buffers and a buffer that is no snapshot
    sz = C.c_size_t()
    assert L.hd_snapshot_size(h, C.byref(sz)) == 0
    sb = C.create_string_buffer(sz.value)
    assert L.hd_snapshot(h, sb, sz.value - 1, C.c_double(0.0)) == INVALID
    assert L.hd_restore(h, sb, sz.value - 1, C.byref(t)) == INVALID
    assert L.hd_restore(h, sb, sz.value, C.byref(t)) == INVALID
    assert L.hd_save_checkpoint(h, None, C.c_double(0.0)) == INVALID
    assert L.hd_save_checkpoint(h, b"/nonexistent-dir/x.ckpt", C.c_double(0.0)) == INVALID
    assert L.hd_load_checkpoint(h, b"/nonexistent-dir/x.ckpt", C.byref(t)) == INVALID
    # constructors and free functions
    assert L.hd_create(None, None, None, None, None) == INVALID
    assert L.hd_batch_create(None, 0, None, None, None, None) == INVALID
    assert L.hd_footprint(None, None, 1, None) == INVALID
    assert L.hd_gtsv_batch(None, None, None, None, None, 4, 1, 4, 0, None) == INVALID
    assert L.hd_server_create(None, None) == INVALID
    assert L.hd_server_poll(None, 1, None, None) == INVALID
    # the context still steps, equal to a fresh one
    g.step(3)
    f = engine.Simulation(sc)
    f.init()
    f.step(3)
    for name in ("U", "V", "Z"):
        assert np.array_equal(g.get(name), f.get(name)), name
    f.close()
    g.close()


def test_server_calls_with_bad_ids_sizes_and_states():
    print("[case] server ABI checks", flush=True)
    L = engine._server_lib()
    sc = S.open_box(20, 14, seed=2)
    srv = engine.Server(hbm_budget=engine.footprint(sc) + 1, max_active=1, steps_per_quantum=1, use_graph=False)
    h = srv.h
    st, tt = C.c_int(), C.c_double()
    assert L.hd_server_poll(h, 12345, C.byref(st), C.byref(tt)) == INVALID
    assert L.hd_server_cancel(h, 12345) == INVALID
    assert L.hd_server_release(h, 12345) == INVALID
    assert L.hd_server_stats(h, C.create_string_buffer(8), 8) == INVALID  # too small for the JSON
    # a mesh below the smallest size is refused at submission
    tiny = engine.MeshHolder(S.open_box(20, 14))
    tiny.mesh.M = 3
    rid = C.c_int64()
    coeffs, ops = engine.load_coeffs(), engine.make_options(sc, 8.0)
    cfg = engine.make_config()
    assert L.hd_server_submit(h, C.byref(tiny.mesh), C.byref(coeffs), C.byref(ops), C.byref(cfg),
                              C.byref(rid)) == INVALID
    # so is a setup hd_create would refuse (here a zero time step)
    good = engine.MeshHolder(sc)
    zero_dt = engine.load_coeffs()
    zero_dt.dT = 0.0
    assert L.hd_server_submit(h, C.byref(good.mesh), C.byref(zero_dt), C.byref(ops), C.byref(cfg),
                              C.byref(rid)) == INVALID
    a = srv.submit(sc, 40000.0)  # holds the one slot until cancelled
    b = srv.submit(S.open_box(20, 14, seed=3), 8.0)  # queued behind it
    out = np.zeros((sc.N + 3, sc.M + 3))
    dp = out.ctypes.data_as(C.POINTER(C.c_double))
    assert L.hd_server_release(h, b) == STATE  # not finished
    assert L.hd_server_result(h, b, abi.F["Z"], dp, out.size) == STATE
    assert srv.wait(b, timeout=0.0) < engine.RUN_DONE
    srv.cancel(b)
    assert srv.wait(b) == engine.RUN_CANCELLED
    assert L.hd_server_result(h, b, abi.F["Z"], dp, out.size) == -8  # HD_E_CANCELLED
    srv.cancel(a)
    assert srv.wait(a) == engine.RUN_CANCELLED
    a = srv.submit(sc, 8.0)
    assert srv.wait(a) == engine.RUN_DONE
    assert L.hd_server_result(h, a, abi.F["Z"], dp, out.size - 1) == INVALID
    assert L.hd_server_result(h, a, abi.F["HMOI"], dp, out.size) == INVALID  # not a result field
    assert L.hd_server_result(h, a, abi.F["Z"], dp, out.size) == 0
    assert L.hd_server_release(h, a) == 0
    assert L.hd_server_poll(h, a, C.byref(st), C.byref(tt)) == INVALID  # released
    srv.close()


def test_create_refuses_bad_meshes_and_setups():
    print("[case] hd_create argument checks", flush=True)
    L = engine.lib()
    sc = S.open_box(20, 14)

    def create(mesh=None, cfg=None, coeffs=None):
        holder = abi.MeshHolder(sc)
        c = engine.load_coeffs()
        o = abi.make_options(sc, 8.0)
        f = abi.make_config()
        for obj, fn in ((holder.mesh, mesh), (f, cfg), (c, coeffs)):
            if fn:
                fn(obj)
        h = C.c_void_p()
        rc = L.hd_create(C.byref(holder.mesh), C.byref(c), C.byref(o), C.byref(f), C.byref(h))
        if rc == 0:
            L.hd_destroy(h)
        return rc

    assert create() == 0
    bad = [
        dict(mesh=lambda m: setattr(m, "M", 3)),
        dict(mesh=lambda m: setattr(m, "N", -5)),
        dict(mesh=lambda m: setattr(m, "h", None)),
        dict(mesh=lambda m: setattr(m, "total_time", 1)),
        dict(mesh=lambda m: m.bc.__setitem__(1, None)),
        dict(cfg=lambda f: setattr(f, "form", 3)),
        dict(cfg=lambda f: setattr(f, "quirks", 0x2)),
        dict(coeffs=lambda c: setattr(c, "dT", 0.0)),
        dict(coeffs=lambda c: setattr(c, "dX", -10.0)),
        dict(coeffs=lambda c: setattr(c, "dY", float("nan"))),
        dict(coeffs=lambda c: setattr(c, "dT", float("inf"))),
    ]
    for kw in bad:
        assert create(**kw) == INVALID, kw
    assert create(cfg=lambda f: setattr(f, "quirks", abi.HD_QUIRK_Q11_LITERAL if hasattr(abi, "HD_QUIRK_Q11_LITERAL")
                                         else 1)) == 0
    assert create(cfg=lambda f: setattr(f, "device", 99)) == -6  # HD_E_NODEV
