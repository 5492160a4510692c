"""Helpers shared by the GPU parity tests: paired GPU / oracle runs from
identical inputs, with the init fields that involve pow() (Kx1, Ky1, hi,
hmax) copied from the GPU into the oracle so both then step from the same
state."""
import numpy as np

import oracle_lib as ol
from hydist_c_amd import engine
from hydist_c_amd.abi import HD_FORM_FP64_PLAIN

HYDRO = ["Z", "U", "V", "TZ", "TU", "TV", "HMOI", "HTDU", "HTDV"]


def bits_equal(a, b):
    """Bitwise equality of float64 arrays.  NaNs must sit at the same
    places; their sign/payload is not compared (x86 and the GPU produce
    default NaNs of opposite sign)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(np.where(na, 0.0, a).view(np.int64), np.where(nb, 0.0, b).view(np.int64))


def ulp_diff(a, b):
    ia = np.ascontiguousarray(a).view(np.int64).astype(np.int64)
    ib = np.ascontiguousarray(b).view(np.int64).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-(2 ** 63)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2 ** 63)) - ib, ib)
    return np.abs(ia - ib)


def make_pair(scen, form=HD_FORM_FP64_PLAIN, sediment=False, bed_change=False, quirks=0, seed_init=True):
    g = engine.Simulation(scen, form=form, sediment=sediment, bed_change=bed_change, quirks=quirks)
    cfg = ol.make_config(sediment=sediment, bed_change=bed_change, quirks=quirks)
    o = ol.Oracle(scen, cfg=cfg)
    g.init()
    o.init()
    if seed_init:
        for f in ("KX1", "KY1"):
            o.set(f, g.get(f))
        o.set_boundary("HI", g.get_boundary("HI"))
        o.set_hmax(g.get_hmax())
    return g, o


def diff_report(g, o, names):
    out = {}
    for n in names:
        a, b = g.get(n), o.get(n)
        fin = np.isfinite(a) & np.isfinite(b)
        out[n] = {
            "bitwise": bits_equal(a, b),
            "max_abs": float(np.max(np.abs(a[fin] - b[fin]))) if fin.any() else 0.0,
            "nan_mismatch": int((np.isnan(a) != np.isnan(b)).sum()),
        }
    return out
