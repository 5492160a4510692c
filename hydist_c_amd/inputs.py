"""Input preparation: the grid transform of Inputs/standardize_inp.py
(lines 4-31) that turns raw maps into the layout load_inputs reads.

Each raw map (rows x cols) is transposed, mirrored along its second axis
and edge-padded by one cell before and two after on both axes; the depth
map also gets NANGDAY added.  So a raw map of R rows and C columns gives a
grid of N + 3 = C + 3 rows and M + 3 = R + 3 columns with
    out[1 + a][1 + b] = raw[R - 1 - b][a].
The reference script rewrites the files in place with numpy's default
text format; `standardize_dir` does the same into a chosen directory.
"""
import os

import numpy as np

# (file, add NANGDAY) for the four maps the reference script transforms
MAPS = (("bandodosau.txt", True), ("hsnham.txt", False), ("hsnhotroiA.txt", False), ("Fw_map.txt", False))


def standardize(raw, nangday=0.0, add_nangday=False):
    """One raw map -> the padded (C + 3) x (R + 3) grid."""
    g = np.flip(np.transpose(np.asarray(raw, dtype=np.float64)), 1)
    g = np.pad(g, ((1, 2), (1, 2)), mode="edge")
    if add_nangday:
        g = g + nangday
    return g


def standardize_dir(src, dst, nangday=0.0):
    """Transforms the four raw maps found in `src` into `dst` (which may be
    `src`, as the reference script does).  Returns the files written."""
    os.makedirs(dst, exist_ok=True)
    out = []
    for name, add in MAPS:
        p = os.path.join(src, name)
        if not os.path.exists(p):
            continue
        g = standardize(np.loadtxt(p), nangday, add)
        q = os.path.join(dst, name)
        np.savetxt(q, g)
        out.append(q)
    return out


def write_scenario_inputs(scen, dirname):
    """The 13 text files load_inputs reads (loader.cpp:120-168) for a
    Scenario, laid out so that a caller in main.cpp's order
    (main.cpp:138: its bc_down sits in load_inputs' bc_right slot) runs
    exactly this scenario: bienphai.txt carries the down side's series and
    bienduoi.txt the right side's.  The grids are already in the padded
    (N+3) x (M+3) layout.  Returns `dirname` with a trailing '/' (the
    reference concatenates file names to it)."""
    os.makedirs(dirname, exist_ok=True)
    W = scen.M + 3

    def grid(name, arr):
        np.savetxt(os.path.join(dirname, name), np.asarray(arr, dtype=np.float64).reshape(-1, W), fmt="%.17g")

    grid("bandodosau.txt", scen.h)
    grid("hsnham.txt", scen.hsnham)
    grid("hsnhotroiA.txt", scen.visc)
    grid("Fw_map.txt", scen.Fw)
    cc = scen.CC if scen.CC is not None else [np.zeros(scen.total_time)] * 4
    series = {"bientren": scen.bc[0], "bienphai": scen.bc[1], "bientrai": scen.bc[2], "bienduoi": scen.bc[3],
              "FS_bientren": cc[0], "FS_bienduoi": cc[1], "FS_bientrai": cc[2], "FS_bienphai": cc[3]}
    for name, v in series.items():
        v = np.asarray(v, dtype=np.float64).reshape(-1)[:scen.total_time]
        np.savetxt(os.path.join(dirname, name + ".txt"), v.reshape(-1, 1), fmt="%.17g")
    np.savetxt(os.path.join(dirname, "boundary_type.txt"), np.asarray(scen.boundary_type).reshape(-1, 1), fmt="%d")
    return dirname.rstrip("/") + "/"
