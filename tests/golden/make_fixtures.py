"""Write the reference-held data files used as golden fixtures.

Reads the shipped real-river inputs and the initial wet/dry mask dump of
the reference checkout (Inputs/bandodosau.txt, Outputs/khouot.txt written
by main.cpp:268-280) and stores them as small compressed .npz files next
to this script.  Run once in a container that has the reference mounted:

    python tests/golden/make_fixtures.py /root/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    inp = os.path.join(ref, "Inputs")
    h = np.loadtxt(os.path.join(inp, "bandodosau.txt"))
    kh = np.loadtxt(os.path.join(ref, "Outputs", "khouot.txt"), dtype=np.int64)
    assert h.shape == kh.shape, (h.shape, kh.shape)
    np.savez_compressed(os.path.join(HERE, "luy_khouot.npz"), h=h, khouot=kh.astype(np.int8))
    # boundary series and types of the shipped case (loader.cpp:148-156)
    series = {}
    for name in ["bientren", "bienduoi", "bientrai", "bienphai",
                 "FS_bientren", "FS_bienduoi", "FS_bientrai", "FS_bienphai"]:
        series[name] = np.loadtxt(os.path.join(inp, name + ".txt")).reshape(-1)
    bt = np.array([int(x) for x in open(os.path.join(inp, "boundary_type.txt")).read().split()])
    hsn = np.loadtxt(os.path.join(inp, "hsnham.txt"))
    fw = np.loadtxt(os.path.join(inp, "Fw_map.txt"))
    visc = np.array(open(os.path.join(inp, "hsnhotroiA.txt")).read().split(), dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "luy_inputs.npz"), hsnham=hsn, Fw=fw,
                        visc_flat=visc, boundary_type=bt, **series)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
