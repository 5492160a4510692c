"""The generators' wet-segment reports: scenarios.segment_lengths (the
reference's segmentation rules, vectorised) equals the oracle's segment
tables after init on every generator family, and the configuration
meshes' histograms are pinned: C2's segments span 1-798 cells (the
configuration asks for 1 to about 900); C3's braided river gives 7 895 row
and 34 541 column segments (its twelve channels of width 10-120 cells)."""
import numpy as np
import pytest

import oracle_lib as ol
from hydist_c_amd import scenarios as S


@pytest.mark.parametrize("make", [
    lambda: S.open_box(57, 33, seed=9, n_islands=5),
    lambda: S.meander_islands(160, 96, n_islands=8),
    lambda: S.tidal_flats(64, 40, seed=6),
    lambda: S.braided(128, 96, seed=5),
    lambda: S.closed_basin(30, 22),
], ids=["box", "islands", "flats", "braided", "basin"])
def test_segment_lengths_match_the_oracle_tables(make):
    sc = make()
    o = ol.Oracle(sc)
    o.init()
    lens = S.segment_lengths(sc.h, sc.M, sc.N)
    for d, key in ((0, "rows"), (1, "cols")):
        counts, first, last = o.segments(d)
        assert sorted((last - first + 1).tolist()) == sorted(lens[key].tolist()), key


def test_config2_histogram():
    sc = S.config2()
    lens = S.segment_lengths(sc.h, sc.M, sc.N)
    hist = S.segment_histogram(sc.h, sc.M, sc.N)
    print({k: v.tolist() for k, v in hist.items()})
    assert (len(lens["rows"]), len(lens["cols"])) == (3712, 3238)
    assert lens["rows"].min() == 1 and lens["rows"].max() == 798
    assert lens["cols"].min() == 1 and lens["cols"].max() == 492
    assert hist["rows"].tolist() == [8, 32, 721, 2951, 0] and hist["cols"].tolist() == [3, 2, 568, 2665, 0]


def test_config3_histogram():
    sc = S.config3()
    lens = S.segment_lengths(sc.h, sc.M, sc.N)
    hist = S.segment_histogram(sc.h, sc.M, sc.N)
    print({k: v.tolist() for k, v in hist.items()})
    assert (len(lens["rows"]), len(lens["cols"])) == (7895, 34541)
    assert lens["rows"].min() == 1 and lens["rows"].max() == 4086
    assert lens["cols"].min() == 10 and lens["cols"].max() == 267
    assert hist["rows"].tolist() == [101, 296, 4029, 2822, 647] and hist["cols"].tolist() == [0, 0, 29838, 4703, 0]
