"""ctypes mirrors of the POD structs in include/hydist.h.

hd_coeffs is Constant_Coeffs (engine.h:39-52) and hd_options the data
members of Options (engine.h:16-37); both keep the reference's member
order and types so the same bytes can be handed to either side.
"""
import ctypes as C

import numpy as np

HD_OK = 0
HD_E_INVALID = -1
HD_E_NOMEM = -2
HD_E_HIP = -3
HD_E_NUMERIC = -4
HD_E_STATE = -5
HD_E_NODEV = -6
HD_E_FAULT = -7
HD_E_CANCELLED = -8
HD_E_UNSUPPORTED = -9

HD_SIDE_UP, HD_SIDE_DOWN, HD_SIDE_LEFT, HD_SIDE_RIGHT = range(4)

HD_FORM_FP64_PLAIN = 0
HD_FORM_FP64_FAST = 1
HD_FORM_FP32_REFINE = 2

HD_QUIRK_Q11_LITERAL = 0x1

FIELDS = [
    "U", "V", "Z", "TU", "TV", "TZ", "HMOI", "HTDU", "HTDV", "FS", "TFS",
    "VTH", "KX", "KY", "QBX", "QBY", "DH", "HTAIZ", "HTAIZ_BD", "KX1", "KY1",
    "H", "HSNHAM", "VISCOIDX", "FW", "KHOUOT",
]
F = {name: k for k, name in enumerate(FIELDS)}
BOUNDARY = {"UBT": 0, "UBP": 1, "VBT": 2, "VBD": 3, "HI": 4}

_COEFF_NAMES = (
    "dX dY dT dXbp dYbp dX2 dY2 dTchia2dX dTchia2dY HaiChiadT "
    "QuyDoiTime QuyDoiPi NANGDAY H_TINH heso mu_mn kenhhepng kenhhepd "
    "Wind huonggio Zbandau NDnen NDbphai NDbtrai NDbtren NDbduoi "
    "Tod Toe hstoe ghtoe Mbochat ro ros dm muy d90 Dorong Sx sohn soha sohn1 "
    "CORIOLIS_FORCE g Ks Windx Windy Ufr Dxr wss"
).split()


class Coeffs(C.Structure):
    _fields_ = [(n, C.c_double) for n in _COEFF_NAMES]

    def as_dict(self):
        return {n: getattr(self, n) for n in _COEFF_NAMES}


class Options(C.Structure):
    _fields_ = [
        ("Tmax", C.c_double), ("t_start", C.c_double),
        ("sediment_start", C.c_double), ("bed_change_start", C.c_double),
        ("interval", C.c_int),
        ("kenhhepd", C.c_bool), ("kenhhepng", C.c_bool),
        ("M", C.c_int), ("N", C.c_int),
        ("ketdinh", C.c_bool), ("channel", C.c_bool), ("debug", C.c_bool), ("plot", C.c_bool),
        ("total_time", C.c_int),
    ]


_dp = C.POINTER(C.c_double)


class Mesh(C.Structure):
    _fields_ = [
        ("M", C.c_int), ("N", C.c_int),
        ("h", _dp), ("hsnham", _dp), ("VISCOIDX", _dp), ("Fw", _dp),
        ("Tsxw", _dp), ("Tsyw", _dp),
        ("u0", _dp), ("v0", _dp), ("z0", _dp), ("FS0", _dp),
        ("total_time", C.c_int),
        ("boundary_type", C.c_int * 4),
        ("bc", _dp * 4),
        ("CC", _dp * 4),
    ]


class Config(C.Structure):
    _fields_ = [
        ("form", C.c_int), ("sediment", C.c_int), ("bed_change", C.c_int),
        ("quirks", C.c_int), ("device", C.c_int), ("dense_layout", C.c_int),
    ]


def dptr(a):
    """Pointer to a C-contiguous float64 array (or NULL for None)."""
    if a is None:
        return C.cast(None, _dp)
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


class MeshHolder:
    """Keeps the numpy arrays behind a Mesh struct alive."""

    def __init__(self, scen):
        self.arrays = []
        m = Mesh()
        m.M, m.N = scen.M, scen.N

        def keep(a):
            if a is None:
                return dptr(None)
            a = np.ascontiguousarray(a, dtype=np.float64)
            self.arrays.append(a)
            return dptr(a)

        m.h = keep(scen.h)
        m.hsnham = keep(scen.hsnham)
        m.VISCOIDX = keep(scen.visc)
        m.Fw = keep(scen.Fw)
        m.Tsxw = keep(getattr(scen, "Tsxw", None))
        m.Tsyw = keep(getattr(scen, "Tsyw", None))
        m.u0 = keep(getattr(scen, "u0", None))
        m.v0 = keep(getattr(scen, "v0", None))
        m.z0 = keep(getattr(scen, "z0", None))
        m.FS0 = keep(getattr(scen, "FS0", None))
        m.total_time = scen.total_time
        for s in range(4):
            m.boundary_type[s] = int(scen.boundary_type[s])
            m.bc[s] = keep(scen.bc[s])
            m.CC[s] = keep(scen.CC[s] if scen.CC is not None else None)
        self.mesh = m


def make_options(scen, Tmax=3600.0):
    o = Options()
    o.Tmax = Tmax
    o.t_start = 0.0
    o.sediment_start = scen.sediment_start
    o.bed_change_start = scen.bed_change_start
    o.interval = 60
    o.kenhhepd = False
    o.kenhhepng = False
    o.M, o.N = scen.M, scen.N
    o.ketdinh = True
    o.channel = False
    o.debug = False
    o.plot = False
    o.total_time = scen.total_time
    for k, v in getattr(scen, "options", {}).items():
        setattr(o, k, v)
    return o


def make_config(form=HD_FORM_FP64_PLAIN, sediment=False, bed_change=False, quirks=0, device=0):
    c = Config()
    c.form, c.sediment, c.bed_change, c.quirks, c.device = form, int(sediment), int(bed_change), quirks, device
    return c
