"""Batched tridiagonal solve: hd_gtsv_batch (the drop-in for
gtsv_spike_partial_diag_pivot_v1, gtsv.cpp:88) and the partition solver
behind it, and hd_gtsv_long (one system of any length, chunked, with an
interface system), against LAPACK's banded solve
(scipy.linalg.solve_banded, with partial pivoting) in FP64.

The systems are diagonally dominant and carry zero couplings and identity
rows, as the lines of the sweeps do.  Rows past m hold NaN and dl[0],
du[m-1] hold NaN: the solver must read none of them.

  FP64:         max|x - x_ref| <= 1e-12 * max|x_ref|
  FP32 + refine: max|x - x_ref| <= 1e-9 * max|x_ref|
"""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.linalg import solve_banded

from hydist_c_amd import engine

pytestmark = pytest.mark.gpu
EMU = os.environ.get("HD_EMU") == "1"


def _hip():
    """The HIP runtime the library links (already loaded by engine.lib())."""
    engine.lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Dev:
    """A float64 device buffer initialised from a host array."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.shape, self.nbytes = host.shape, host.nbytes
        if EMU:
            L = engine.lib()
            L.hd_emu_alloc.restype = C.c_void_p
            L.hd_emu_alloc.argtypes = [C.c_size_t]
            self.ptr = L.hd_emu_alloc(max(self.nbytes, 8))
            C.memmove(self.ptr, host.ctypes.data, self.nbytes)
        else:
            p = C.c_void_p()
            assert _hip().hipMalloc(C.byref(p), max(self.nbytes, 8)) == 0
            self.ptr = p.value
            assert _hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        if EMU:
            C.memmove(out.ctypes.data, self.ptr, self.nbytes)
        else:
            assert _hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if EMU:
            L = engine.lib()
            L.hd_emu_free.argtypes = [C.c_void_p]
            L.hd_emu_free(self.ptr)
        else:
            _hip().hipFree(self.ptr)


def make_systems(rng, m, nsys, ld):
    dl = rng.uniform(-1, 1, (nsys, ld))
    du = rng.uniform(-1, 1, (nsys, ld))
    d = (np.abs(dl) + np.abs(du) + rng.uniform(0.05, 2.0, (nsys, ld))) * rng.choice([-1.0, 1.0], (nsys, ld))
    b = rng.uniform(-1, 1, (nsys, ld))
    # zero couplings (segment ends) and identity rows (dry cells)
    cut = rng.random((nsys, ld)) < 0.05
    du[cut] = 0.0
    dl[:, 1:][cut[:, :-1]] = 0.0
    ident = rng.random((nsys, ld)) < 0.03
    dl[ident], du[ident], d[ident], b[ident] = 0.0, 0.0, 1.0, 0.0
    du[:, :-1][ident[:, 1:]] = 0.0
    dl[:, 1:][ident[:, :-1]] = 0.0
    for a in (dl, d, du, b):
        a[:, m:] = np.nan
    dl[:, 0] = np.nan
    du[:, m - 1] = np.nan
    return dl, d, du, b


def reference(dl, d, du, b, m):
    x = np.empty((dl.shape[0], m))
    for s in range(dl.shape[0]):
        ab = np.zeros((3, m))
        ab[0, 1:] = du[s, :m - 1]
        ab[1] = d[s, :m]
        ab[2, :-1] = dl[s, 1:m]
        x[s] = solve_banded((1, 1), ab, b[s, :m])
    return x


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32r"])
@pytest.mark.parametrize("m", [1, 2, 3, 7, 64, 65, 300, 1023, 2048, 2049, 4096])
def test_gtsv_batch_matches_lapack(m, precision):
    print(f"[case] gtsv m={m} precision={precision}", flush=True)
    rng = np.random.default_rng(1000 + m)
    nsys, ld = 3, m + 5
    dl, d, du, b = make_systems(rng, m, nsys, ld)
    ref = reference(dl, d, du, b, m)
    bufs = [Dev(a) for a in (dl, d, du, b)]
    xb = Dev(np.full((nsys, ld), 7.0))
    engine.gtsv_batch(*[q.ptr for q in bufs], xb.ptr, m, nsys, ld, precision)
    x = xb.get()
    for q in bufs + [xb]:
        q.free()
    assert np.all(x[:, m:] == 7.0), "wrote past m"
    err = np.max(np.abs(x[:, :m] - ref))
    tol = (1e-12 if precision == 0 else 1e-9) * np.max(np.abs(ref))
    print(f"[gtsv] m={m} p={precision} max err {err:.3e} (tol {tol:.1e})", flush=True)
    assert err <= tol, (m, precision, err)


def test_gtsv_batch_rejects_bad_shapes():
    print("[case] gtsv shapes", flush=True)
    one = Dev(np.zeros(8))
    p = one.ptr
    for m, nsys, ld, prec in ((0, 1, 1, 0), (4097, 1, 4097, 0), (8, 2, 4, 0), (8, 1, 8, 2)):
        with pytest.raises(engine.HydistError):
            engine.gtsv_batch(p, p, p, p, p, m, nsys, ld, prec)
    engine.gtsv_batch(p, p, p, p, p, 8, 0, 8, 0)  # no systems: nothing to do
    one.free()


LONG_M = [4097, 5000, 8192, 8193, 12289] + ([] if EMU else [100_003, 1_000_000])


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32r"])
@pytest.mark.parametrize("m", LONG_M)
def test_gtsv_long_matches_lapack(m, precision):
    """Lengths past one chunk (4096 rows): a remainder of one row (8193),
    of many (5000, 100003), none (8192) and a single-row last chunk
    (12289 = 3 * 4096 + 1)."""
    print(f"[case] gtsv_long m={m} precision={precision}", flush=True)
    rng = np.random.default_rng(7000 + m)
    dl, d, du, b = make_systems(rng, m, 1, m + 5)
    ref = reference(dl, d, du, b, m)[0]
    bufs = [Dev(a[0]) for a in (dl, d, du, b)]
    xb = Dev(np.full(m + 5, 7.0))
    engine.gtsv_long(*[q.ptr for q in bufs], xb.ptr, m, precision)
    x = xb.get()
    for q in bufs + [xb]:
        q.free()
    assert np.all(x[m:] == 7.0), "wrote past m"
    err = np.max(np.abs(x[:m] - ref))
    tol = (1e-12 if precision == 0 else 1e-9) * np.max(np.abs(ref))
    print(f"[gtsv_long] m={m} p={precision} max err {err:.3e} (tol {tol:.1e})", flush=True)
    assert err <= tol, (m, precision, err)


def test_gtsv_long_rejects_bad_shapes():
    print("[case] gtsv_long shapes", flush=True)
    one = Dev(np.zeros(8))
    p = one.ptr
    for m, prec in ((0, 0), (4096 * 4096 + 1, 0), (8, 2)):
        with pytest.raises(engine.HydistError):
            engine.gtsv_long(p, p, p, p, p, m, prec)
    one.free()
