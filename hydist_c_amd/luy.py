"""The reference's shipped real-river case (Luy river, 464 x 324) as a
Scenario, from the golden fixtures (tests/golden/luy_*.npz), and as the
13 input text files load_inputs reads (loader.cpp:120-168)."""
import os

import numpy as np

from .scenarios import Scenario

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def load_fixture():
    a = np.load(os.path.join(GOLDEN, "luy_khouot.npz"))
    b = np.load(os.path.join(GOLDEN, "luy_inputs.npz"))
    return a, b


def scenario(swap_down_right=True):
    """Scenario with the reference's main.cpp semantics: main.cpp:138 passes
    its bc_down in load_inputs' bc_right slot, so the down side receives
    bienphai.txt and the right side bienduoi.txt (swap_down_right)."""
    a, b = load_fixture()
    h = np.ascontiguousarray(a["h"])
    H, W = h.shape
    N, M = H - 3, W - 3
    n = H * W
    visc = np.ascontiguousarray(b["visc_flat"][:n].reshape(H, W))  # flat read, loader.cpp:91-112
    bt = list(b["boundary_type"]) + [0] * 4
    total_time = len(b["bienduoi"])
    down, right = (b["bienphai"], b["bienduoi"]) if swap_down_right else (b["bienduoi"], b["bienphai"])
    fs_down, fs_right = b["FS_bienduoi"], b["FS_bienphai"]

    def ser(x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        out = np.empty(total_time)
        k = min(total_time, x.size)
        out[:k] = x[:k]
        out[k:] = x[k - 1]
        return out

    return Scenario(
        name="luy_464x324", M=M, N=N, h=h, hsnham=np.ascontiguousarray(b["hsnham"]), visc=visc,
        Fw=np.ascontiguousarray(b["Fw"]), total_time=total_time, boundary_type=tuple(int(x) for x in bt[:4]),
        bc=[ser(b["bientren"]), ser(down), ser(b["bientrai"]), ser(right)],
        CC=[ser(b["FS_bientren"]), ser(fs_down), ser(b["FS_bientrai"]), ser(fs_right)],
        sediment_start=36000.0, bed_change_start=1.8e6,
    )


def write_inputs(dirname):
    """The 13 text files of the shipped case, as load_inputs reads them."""
    a, b = load_fixture()
    os.makedirs(dirname, exist_ok=True)
    W = a["h"].shape[1]

    def grid(name, arr, cols):
        arr = np.asarray(arr, dtype=np.float64).reshape(-1, cols)
        np.savetxt(os.path.join(dirname, name), arr, fmt="%.17g")

    grid("bandodosau.txt", a["h"], W)
    grid("hsnham.txt", b["hsnham"], W)
    grid("Fw_map.txt", b["Fw"], W)
    grid("hsnhotroiA.txt", b["visc_flat"], 338)
    for name in ("bientren", "bienduoi", "bientrai", "bienphai",
                 "FS_bientren", "FS_bienduoi", "FS_bientrai", "FS_bienphai"):
        np.savetxt(os.path.join(dirname, name + ".txt"), np.asarray(b[name]).reshape(-1, 1), fmt="%.17g")
    np.savetxt(os.path.join(dirname, "boundary_type.txt"), np.asarray(b["boundary_type"]).reshape(-1, 1),
               fmt="%d")
    return dirname
