"""GPU forms against the CPU oracle on degenerate wet/dry layouts.

The meshes stress the segment rules (support_funcs.cpp:238-343) and the
end-type rows of the assembly (UVZSolver_multithread.cpp:110-139, 345-470)
where the ordinary scenarios rarely go:

  speckled  random wet cells: hundreds of segments per sweep, a third to
            a half of them one or two cells long (a one-cell segment is a
            one-row hydraulic system, sn == 0, or no FS system at all)
  stripes   two land rows in every four: each column is cut into
            two-cell runs
  onecell   a single 2 x 2 patch of water in dry land
  alland    no wet cell: no segment in either direction

Plain FP64: hydraulic fields bit for bit, mask and segment tables equal,
FS and dH to the tolerance of the sediment parity test (device pow/log).
Fast forms: the tolerances of the fast-form tests.  Every case runs
sediment and bed change from t = 0.

The fast forms are compared on `stripes` for one step only.  Each column
there is a closed two-cell channel, so the along-column velocity u is
zero up to round-off (a few 1e-18).  The VZ preindex switches its upwind
terms on the sign of the average of four such values next to dry cells
(`if (utb < 0)`, UVZSolver_multithread.cpp:186-191): the plain form
reproduces the oracle's round-off and takes the same branch, but a
different elimination order leaves noise of the other sign in the second
step, the curvature term turns on in one run and off in the other, and Z
moves by about 5e-4.  That is the reference scheme's own sensitivity,
not a solver error; the first step is checked to the usual bar."""
import numpy as np
import pytest

from hydist_c_amd import scenarios as S
from hydist_c_amd.abi import HD_FORM_FP32_REFINE, HD_FORM_FP64_FAST
from parity_util import HYDRO, diff_report, make_pair

pytestmark = pytest.mark.gpu

STEPS = 20
FAST_STEPS = {"stripes": 1}  # see the module docstring
# hydraulics (m, m/s) and FS / bed change relative to the field's largest
# value: about 100 times the largest error observed on these cases (fast
# FP64 ~1e-15, FP32 + refine a few 1e-12 on the hydraulics)
TOL = {HD_FORM_FP64_FAST: 1e-13, HD_FORM_FP32_REFINE: 5e-10}
REL_TOL = {HD_FORM_FP64_FAST: 1e-12, HD_FORM_FP32_REFINE: 1e-9}
NAMES = {HD_FORM_FP64_FAST: "fp64_fast", HD_FORM_FP32_REFINE: "fp32_refine"}


def _from_mask(M, N, in_channel, name, seed=3):
    """An open box (Q inflow on two sides, water levels on the other two)
    whose wet cells are exactly `in_channel(y, x)` (cell centres in core
    coordinates), 3 m deep."""
    sc = S.open_box(M, N, seed=seed, n_islands=0)
    sc.h = S.build_bathymetry(M, N, in_channel, lambda y, x: np.full(np.broadcast(y, x).shape, 3.0))
    sc.name = name
    return sc


def speckled(M, N, p_wet, seed):
    rng = np.random.default_rng(seed)
    wet = rng.random((N + 3, M + 3)) < p_wet

    def in_channel(y, x):
        i = np.clip((y + 1.5).astype(int), 0, N + 2)
        j = np.clip((x + 1.5).astype(int), 0, M + 2)
        return wet[i, j]

    return _from_mask(M, N, in_channel, f"speckled{int(p_wet * 100)}_{M}x{N}", seed=seed)


MESHES = {
    "speckled25": lambda: speckled(48, 32, 0.25, 11),
    "speckled35": lambda: speckled(48, 32, 0.35, 11),
    "stripes": lambda: _from_mask(40, 24, lambda y, x: np.floor(y).astype(int) % 4 >= 2, "stripes_40x24"),
    "onecell": lambda: _from_mask(40, 24, lambda y, x: (np.abs(y - 10) < 0.6) & (np.abs(x - 12) < 0.6),
                                  "onecell_40x24"),
    "alland": lambda: _from_mask(40, 24, lambda y, x: np.zeros(np.broadcast(y, x).shape, bool), "alland_40x24"),
}


def _segment_lengths(o):
    lens = []
    for d in (0, 1):
        _, f, l = o.segments(d)
        lens.extend(list(l - f + 1))
    return np.array(lens, dtype=np.int64)


def _check_layout(name, o):
    """The mesh has the shape its name promises (guards the generator)."""
    lens = _segment_lengths(o)
    if name.startswith("speckled"):
        assert lens.size >= 300 and (lens == 1).sum() >= 100 and (lens == 2).sum() >= 40, lens.size
    elif name == "stripes":
        assert lens.size > 0 and (lens == 2).sum() >= 20
    elif name == "onecell":
        assert 0 < lens.size <= 4 and lens.max() <= 2
    else:
        assert lens.size == 0 and not (o.get("KHOUOT") == 0).any()


@pytest.mark.parametrize("mesh", list(MESHES))
def test_edge_mesh_plain_fp64_bitwise(mesh):
    sc = MESHES[mesh]()
    print(f"[case] edge-plain {mesh} {sc.name}", flush=True)
    g, o = make_pair(sc, sediment=True, bed_change=True)
    _check_layout(mesh, o)
    g.step(STEPS)
    o.step(STEPS)
    rep = diff_report(g, o, HYDRO)
    bad = {k: v for k, v in rep.items() if not v["bitwise"]}
    assert not bad, bad
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for d in (0, 1):
        assert all(np.array_equal(a, b) for a, b in zip(g.segments(d), o.segments(d)))
    for f, rtol in (("FS", 1e-10), ("DH", 1e-9)):
        a, b = g.get(f), o.get(f)
        scale = np.abs(b).max() + 1e-300
        assert np.max(np.abs(a - b)) <= rtol * scale, (f, np.max(np.abs(a - b)), scale)
    if mesh == "alland":
        for f in ("Z", "U", "V"):
            assert not np.any(g.get(f)), f


@pytest.mark.parametrize("form", [HD_FORM_FP64_FAST, HD_FORM_FP32_REFINE], ids=NAMES.get)
@pytest.mark.parametrize("mesh", list(MESHES))
def test_edge_mesh_fast_forms(form, mesh):
    sc = MESHES[mesh]()
    print(f"[case] edge-fast {mesh} {NAMES[form]} {sc.name}", flush=True)
    g, o = make_pair(sc, form=form, sediment=True, bed_change=True)
    steps = FAST_STEPS.get(mesh, STEPS)
    g.step(steps)
    o.step(steps)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for f in ("Z", "U", "V"):
        d = np.max(np.abs(g.get(f) - o.get(f)))
        print(f"[{f}] max abs error {d:.3e}", flush=True)
        assert d <= TOL[form], (f, d)
    for f in ("FS", "DH"):
        a, b = g.get(f), o.get(f)
        scale = np.abs(b).max() + 1e-300
        print(f"[{f}] relative error {np.max(np.abs(a - b)) / scale:.3e}", flush=True)
        assert np.max(np.abs(a - b)) <= REL_TOL[form] * scale, (f, np.max(np.abs(a - b)), scale)


def test_segments_of_a_wide_speckled_mesh():
    """Segment tables after init on a mesh wide and tall enough that the
    segmentation kernel takes several load batches per row and column
    (rows of 1300 positions, columns of 700), equal to the oracle's."""
    sc = speckled(1300, 700, 0.3, 21)
    print(f"[case] edge-segments {sc.name}", flush=True)
    g, o = make_pair(sc, seed_init=False)
    assert np.array_equal(g.get("KHOUOT"), o.get("KHOUOT"))
    for d in (0, 1):
        cg, fg, lg = g.segments(d)
        co, fo, lo = o.segments(d)
        assert fo.size > 10000
        assert np.array_equal(cg, co) and np.array_equal(fg, fo) and np.array_equal(lg, lo)

