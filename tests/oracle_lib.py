"""ctypes binding of the CPU oracle (oracle/_build/libhydist_oracle.so).

Test infrastructure: the oracle is the checker of the GPU library and is
never loaded by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from hydist_c_amd import abi
from hydist_c_amd.abi import F, MeshHolder, make_config, make_options

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libhydist_oracle.so")

NATIVE_PATH = os.path.join(ROOT, "oracle", "_build", "native", "libhydist_oracle.so")
SANITIZED_PATH = os.path.join(ROOT, "oracle", "_build", "sanitize", "libhydist_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def cpu_model():
    with open("/proc/cpuinfo") as f:
        for line in f:
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    return "unknown"


def use_sanitized_build():
    """Switches this process to the AddressSanitizer + UBSan copy of the
    oracle (`make sanitize`); the process must have libasan preloaded."""
    global LIB_PATH, _lib
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    LIB_PATH = SANITIZED_PATH
    _lib = None


def use_native_build():
    """Switches this process to the -O3 -march=native copy of the oracle
    (timing only: it contracts a*b+c, so its bits differ from the
    checker's), building it first when it is missing or was built on
    another CPU.  Returns the CPU model it was built for."""
    global LIB_PATH, _lib
    host = os.path.join(os.path.dirname(NATIVE_PATH), "host.txt")
    model = cpu_model()
    built_for = open(host).readline().strip() if os.path.exists(host) else None
    if not os.path.exists(NATIVE_PATH) or built_for != model:
        subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native"], check=True)
    LIB_PATH = NATIVE_PATH
    _lib = None
    return model


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.ho_create.argtypes = [C.POINTER(abi.Mesh), C.POINTER(abi.Coeffs), C.POINTER(abi.Options),
                                C.POINTER(abi.Config), C.POINTER(vp)]
        for name in ["ho_init", "ho_destroy"]:
            getattr(L, name).argtypes = [vp]
        for name in ["ho_update_boundary", "ho_halfstep_uz", "ho_halfstep_vz", "ho_sediment_j", "ho_sediment_i",
                     "ho_bedload"]:
            getattr(L, name).argtypes = [vp, C.c_double]
        L.ho_step.argtypes = [vp, dp]
        L.ho_run_stage.argtypes = [vp, C.c_char_p]
        L.ho_run.argtypes = [vp, dp, C.c_double]
        for name in ["ho_copy_out", "ho_copy_in", "ho_copy_out_boundary", "ho_copy_in_boundary"]:
            getattr(L, name).argtypes = [vp, C.c_int, dp, C.c_size_t]
        for name in ["ho_copy_out_i32", "ho_copy_in_i32"]:
            getattr(L, name).argtypes = [vp, C.c_int, ip, C.c_size_t]
        L.ho_get_hmax.argtypes = [vp, dp]
        L.ho_set_hmax.argtypes = [vp, dp]
        L.ho_get_segments.argtypes = [vp, C.c_int, ip, C.c_size_t, ip, ip, C.c_size_t,
                                      C.POINTER(C.c_int64)]
        L.ho_mass_balance.argtypes = [vp, dp, dp]
        L.ho_last_error.argtypes = [vp]
        L.ho_last_error.restype = C.c_char_p
        L.ho_thomas.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.ho_segment_line.argtypes = [ip, C.c_int, ip, ip, C.c_int]
        L.ho_line_system.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, C.c_int]
        L.ho_load_coeffs.argtypes = [C.POINTER(abi.Coeffs)]
        L.ho_finalize_coeffs.argtypes = [C.POINTER(abi.Coeffs)]
        L.ho_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def load_coeffs(overrides=None):
    c = abi.Coeffs()
    assert lib().ho_load_coeffs(C.byref(c)) == 0
    if overrides:
        for k, v in overrides.items():
            setattr(c, k, v)
        assert lib().ho_finalize_coeffs(C.byref(c)) == 0
    return c


def set_threads(n):
    lib().ho_set_threads(int(n))


class Oracle:
    """One oracle scenario; the method names follow the product binding."""

    def __init__(self, scen, cfg=None, Tmax=3600.0, coeffs=None):
        self.scen = scen
        self.shape = scen.shape
        self.holder = MeshHolder(scen)
        self.coeffs = coeffs if coeffs is not None else load_coeffs(scen.coeff_overrides)
        self.ops = make_options(scen, Tmax)
        self.cfg = cfg if cfg is not None else make_config()
        h = C.c_void_p()
        rc = lib().ho_create(C.byref(self.holder.mesh), C.byref(self.coeffs), C.byref(self.ops),
                             C.byref(self.cfg), C.byref(h))
        assert rc == 0, rc
        self.h = h
        self.t = 0.0

    def __del__(self):
        if getattr(self, "h", None) is not None:
            lib().ho_destroy(self.h)
            self.h = None

    def init(self):
        assert lib().ho_init(self.h) == 0

    def step(self, n=1):
        t = C.c_double(self.t)
        for _ in range(n):
            rc = lib().ho_step(self.h, C.byref(t))
            assert rc == 0, rc
        self.t = t.value

    def halfstep_uz(self, t):
        assert lib().ho_halfstep_uz(self.h, t) == 0

    def kernel(self, name):
        """One kernel restatement on the current state (ho_run_stage)."""
        assert lib().ho_run_stage(self.h, name.encode()) == 0, name

    def stage(self, name, t):
        """Stage block: 'sediment_j', 'sediment_i' or 'bedload'."""
        assert getattr(lib(), "ho_" + name)(self.h, t) == 0

    def halfstep_vz(self, t):
        assert lib().ho_halfstep_vz(self.h, t) == 0

    def get(self, name):
        if name == "KHOUOT":
            a = np.empty(self.shape, dtype=np.int32)
            assert lib().ho_copy_out_i32(self.h, F[name], a.ctypes.data_as(C.POINTER(C.c_int32)), a.size) == 0
            return a
        a = np.empty(self.shape, dtype=np.float64)
        assert lib().ho_copy_out(self.h, F[name], _dp(a), a.size) == 0
        return a

    def set(self, name, a):
        if name == "KHOUOT":
            a = np.ascontiguousarray(a, dtype=np.int32)
            assert lib().ho_copy_in_i32(self.h, F[name], a.ctypes.data_as(C.POINTER(C.c_int32)), a.size) == 0
            return
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert lib().ho_copy_in(self.h, F[name], _dp(a), a.size) == 0

    def boundary_size(self, which):
        M, N = self.scen.M, self.scen.N
        return {"UBT": M + 3, "UBP": M + 3, "VBT": N + 3, "VBD": N + 3, "HI": 2 * (M + N + 6)}[which]

    def get_boundary(self, which):
        a = np.empty(self.boundary_size(which))
        assert lib().ho_copy_out_boundary(self.h, abi.BOUNDARY[which], _dp(a), a.size) == 0
        return a

    def set_boundary(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert lib().ho_copy_in_boundary(self.h, abi.BOUNDARY[which], _dp(a), a.size) == 0

    def get_hmax(self):
        a = np.empty(4)
        lib().ho_get_hmax(self.h, _dp(a))
        return a

    def set_hmax(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        lib().ho_set_hmax(self.h, _dp(a))

    def segments(self, direction):
        nlines = self.shape[0] if direction == 0 else self.shape[1]
        counts = np.zeros(nlines, dtype=np.int32)
        total = C.c_int64()
        lib().ho_get_segments(self.h, direction, counts.ctypes.data_as(C.POINTER(C.c_int32)), nlines,
                              None, None, 0, C.byref(total))
        n = total.value
        first = np.zeros(max(n, 1), dtype=np.int32)
        last = np.zeros(max(n, 1), dtype=np.int32)
        lib().ho_get_segments(self.h, direction, counts.ctypes.data_as(C.POINTER(C.c_int32)), nlines,
                              first.ctypes.data_as(C.POINTER(C.c_int32)),
                              last.ctypes.data_as(C.POINTER(C.c_int32)), n, C.byref(total))
        return counts, first[:n], last[:n]

    def mass_balance(self):
        v, q = C.c_double(), C.c_double()
        lib().ho_mass_balance(self.h, C.byref(v), C.byref(q))
        return v.value, q.value

    def line_coeffs(self, direction, line, seg):
        """{f1, f2, f3, f5, a2, c2, d2: array over line positions} of one
        hydraulic segment (ho_line_coeffs) and the segment's last position."""
        L = lib()
        L.ho_line_coeffs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int]
        cap = max(self.scen.M, self.scen.N) + 8
        out = np.zeros(7 * cap)
        last = L.ho_line_coeffs(self.h, direction, line, seg, _dp(out), cap)
        assert last >= 0, last
        out = out.reshape(7, cap)
        return {k: out[a] for a, k in enumerate(("f1", "f2", "f3", "f5", "a2", "c2", "d2"))}, last

    def fs_line_system(self, direction, line, seg, t, s_start):
        """(sn, [AA, BB, CC, DD]) of one FS segment (ho_fs_line_system)."""
        L = lib()
        L.ho_fs_line_system.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double] + \
            [C.POINTER(C.c_double)] * 4 + [C.c_int]
        cap = max(self.scen.M, self.scen.N) + 8
        arrs = [np.zeros(cap) for _ in range(4)]
        sn = L.ho_fs_line_system(self.h, direction, line, seg, t, s_start, *[_dp(a) for a in arrs], cap)
        assert sn >= -1, sn
        return sn, [a[:sn + 1] for a in arrs]

    def line_system(self, direction, line, seg):
        cap = 2 * max(self.scen.M, self.scen.N) + 8
        arrs = [np.zeros(cap) for _ in range(5)]
        sn = lib().ho_line_system(self.h, direction, line, seg, *[_dp(a) for a in arrs], cap)
        assert sn >= -1, sn
        return sn, [a[:sn + 1] for a in arrs]


def thomas(A, B, Cc, D):
    n = len(B)
    x = np.zeros(n)
    Ap = np.zeros(n)
    Bp = np.zeros(n)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (A, B, Cc, D)]
    assert lib().ho_thomas(n - 1, *[_dp(a) for a in arrs], _dp(x), _dp(Ap), _dp(Bp)) == 0
    return x


def segment_line(line, end):
    line = np.ascontiguousarray(line, dtype=np.int32)
    cap = len(line)
    f = np.zeros(cap, dtype=np.int32)
    l = np.zeros(cap, dtype=np.int32)
    n = lib().ho_segment_line(line.ctypes.data_as(C.POINTER(C.c_int32)), end,
                              f.ctypes.data_as(C.POINTER(C.c_int32)), l.ctypes.data_as(C.POINTER(C.c_int32)), cap)
    return list(zip(f[:n].tolist(), l[:n].tolist()))


def source_chs(coeffs, wsm, Zf, fs, fw, vth, h):
    L = lib()
    L.ho_source_chs.argtypes = [C.POINTER(abi.Coeffs)] + [C.c_double] * 6
    L.ho_source_chs.restype = C.c_double
    return L.ho_source_chs(C.byref(coeffs), wsm, Zf, fs, fw, vth, h)


def source_nchs(coeffs, wsm, Zf, Uf, h, fs):
    L = lib()
    L.ho_source_nchs.argtypes = [C.POINTER(abi.Coeffs)] + [C.c_double] * 5
    L.ho_source_nchs.restype = C.c_double
    return L.ho_source_nchs(C.byref(coeffs), wsm, Zf, Uf, h, fs)
