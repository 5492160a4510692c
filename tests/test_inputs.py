"""The raw-map transform of Inputs/standardize_inp.py:4-31."""
import numpy as np

from hydist_c_amd import inputs


def test_standardize_index_map_and_padding():
    rng = np.random.default_rng(0)
    R, C = 7, 5
    raw = rng.normal(size=(R, C))
    g = inputs.standardize(raw)
    assert g.shape == (C + 3, R + 3)
    for a in range(C):
        for b in range(R):
            assert g[1 + a, 1 + b] == raw[R - 1 - b, a]
    # edge padding: one row/column before, two after, copies of the border
    assert np.array_equal(g[0, 1:-2], g[1, 1:-2])
    assert np.array_equal(g[-1, 1:-2], g[-3, 1:-2]) and np.array_equal(g[-2, 1:-2], g[-3, 1:-2])
    assert np.array_equal(g[:, 0], g[:, 1]) and np.array_equal(g[:, -1], g[:, -3])


def test_nangday_only_on_depth(tmp_path):
    raw = np.arange(12.0).reshape(3, 4)
    for name, _ in inputs.MAPS:
        np.savetxt(tmp_path / name, raw)
    out = tmp_path / "std"
    files = inputs.standardize_dir(str(tmp_path), str(out), nangday=1.5)
    assert len(files) == 4
    h = np.loadtxt(out / "bandodosau.txt")
    n = np.loadtxt(out / "hsnham.txt")
    assert np.allclose(h - n, 1.5)
    assert np.array_equal(n, inputs.standardize(raw))
