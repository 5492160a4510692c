"""The reference's host surface (include/hydist_compat.h) on the shipped
real-river case: a driver written against load_inputs /
attribute_arrays_memory_alloc / Hydraulic_Calculation / save_result runs
five simulated minutes, and its text output matches the C-ABI path and the
oracle on the same inputs (to the 6 significant digits save_result
writes, loader.cpp:384-388)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
from hydist_c_amd import engine, luy
from parity_util import HYDRO, diff_report, make_pair

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.environ.get("HD_EMU") == "1"
DRIVER = os.path.join(ROOT, "tests", "_build", "emu" if EMU else "", "compat_driver")


def read_grid(path):
    return np.loadtxt(path)


def test_compat_driver_real_river(tmp_path):
    inputs = luy.write_inputs(str(tmp_path / "Inputs"))
    for d in ("U", "V", "Z", "FS"):
        os.makedirs(tmp_path / "Outputs" / d, exist_ok=True)
    print("[case] compat driver, Luy river, 5 min", flush=True)
    r = subprocess.run([DRIVER, "-f", inputs + "/", "-h", "0", "-m", "5", "-s", "0", "-i", "60"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=1500 if EMU else 240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    z_drv = read_grid(tmp_path / "Outputs" / "Z" / "z_300.txt")
    u_drv = read_grid(tmp_path / "Outputs" / "U" / "u_300.txt")
    # save_result after the VZ half-step runs before update_uvz
    # (engine.cpp:243-248): z_300 holds the state after the UZ half of t=299
    sc = luy.scenario()
    g, o = make_pair(sc)
    g.step(149)
    o.step(149)
    g.halfstep_uz(299.0)
    o.halfstep_uz(299.0)
    rep = diff_report(g, o, HYDRO)
    assert all(v["bitwise"] for v in rep.values()), rep
    z = g.get("Z")[1:sc.N + 1, 1:sc.M + 1]
    u = g.get("U")[1:sc.N + 1, 1:sc.M + 1]
    assert z_drv.shape == z.shape
    assert np.allclose(z_drv, z, rtol=1e-5, atol=1e-12)
    assert np.allclose(u_drv, u, rtol=1e-5, atol=1e-12)
    assert os.path.exists(tmp_path / "Outputs" / "Z" / "z_60.txt")
