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
