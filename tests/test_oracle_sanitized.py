"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host
code only): a child process with libasan preloaded loads the sanitized
build (`make -C oracle sanitize`) and runs every stage the checker uses
(init, both half-steps, the FS sweeps and bed change, drying and
re-wetting, segmentation of a speckled mesh, the tridiagonal solve).  Any
out-of-bounds access or undefined behaviour aborts the child."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SCRIPT = r"""
import sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import numpy as np
import oracle_lib as ol
from hydist_c_amd import scenarios as S
ol.use_sanitized_build()
for sc, steps in ((S.open_box(40, 28, seed=3), 6), (S.tidal_flats(48, 32, seed=6), 6),
                  (S.meander_islands(96, 64, n_islands=4), 4)):
    sc.sediment_start = sc.bed_change_start = 0.0
    o = ol.Oracle(sc, cfg=ol.make_config(sediment=True, bed_change=True))
    o.init()
    o.step(steps)
    assert np.isfinite(o.get("Z")).all()
# one- and two-cell runs everywhere: segmentation and the sweeps on a
# speckled mesh (a third of the cells wet at random)
from test_gpu_edge_meshes import MESHES
for name in ("speckled35", "stripes", "onecell", "alland"):
    sc = MESHES[name]()
    sc.sediment_start = sc.bed_change_start = 0.0
    o = ol.Oracle(sc, cfg=ol.make_config(sediment=True, bed_change=True))
    o.init()
    o.step(3)
    for d in (0, 1):
        o.segments(d)
print("sanitized oracle ok", flush=True)
"""


def test_oracle_stages_under_asan_and_ubsan():
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(asan) and os.path.exists(asan), asan
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT, tests=HERE)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0 and "sanitized oracle ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
