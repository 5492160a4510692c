"""The re-wetting pass visits only marked rows (t_z > z with a dry
neighbour at the start of the scan) and follows each downward cascade;
the reference scans every row of the column in order
(support_funcs.cpp:160-175).  This checks on random columns, many of them
mostly dry, that both orders leave identical t_z, khouot and FS."""
import numpy as np


def scan_all(tz, z, kh, fs):
    n = len(tz) - 2  # rows 1..n+1 hold data; the scan covers 2..N, N = n
    for i in range(2, n + 1):
        _row(i, tz, z, kh, fs)


def _row(i, tz, z, kh, fs):
    down = False
    if tz[i] > z[i]:
        if kh[i - 1] == 1:
            tz[i - 1] = tz[i]
            kh[i - 1] = 0
            fs[i - 1] = fs[i]
        if kh[i + 1] == 1:
            tz[i + 1] = tz[i]
            kh[i + 1] = 0
            fs[i + 1] = fs[i]
            down = True
    return down


def scan_marked(tz, z, kh, fs):
    n = len(tz) - 2
    marked = [i for i in range(2, n + 1) if tz[i] > z[i] and (kh[i - 1] == 1 or kh[i + 1] == 1)]
    nxt = 2
    for i in marked:
        if i < nxt:
            continue
        r = i
        while _row(r, tz, z, kh, fs) and r < n:
            r += 1
        nxt = r + 1


def test_marked_scan_equals_full_scan():
    rng = np.random.default_rng(7)
    for trial in range(3000):
        n = int(rng.integers(3, 40))
        rows = n + 2
        p_dry = rng.uniform(0.05, 0.95)
        kh = np.where(rng.random(rows) < p_dry, 1, 0).astype(int)
        kh[rng.random(rows) < 0.05] = 2
        z = rng.normal(size=rows).round(1)
        tz = z + rng.choice([-0.2, 0.0, 0.3], size=rows, p=[0.3, 0.3, 0.4])
        fs = rng.random(rows)
        a = (tz.copy(), z, kh.copy(), fs.copy())
        b = (tz.copy(), z, kh.copy(), fs.copy())
        scan_all(*a)
        scan_marked(*b)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (trial, n)
