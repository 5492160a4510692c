"""Every form gives the same bits whatever order the workgroups of a launch
and the lanes between barriers run in: the library's kernels, run on the
CPU emulator (tests/emu) in launch order, in reverse order
(HD_EMU_ORDER=reverse) and in a pseudo-random order drawn anew for every
launch and pass between barriers (HD_EMU_ORDER=shuffle:SEED), in child
processes, end with identical fields.
The only atomics are bit-ors and the reductions run in a fixed order, so
any difference is a race (a missing barrier, or a workgroup reading what
another writes in the same launch), which on the GPU would make results
vary from run to run.  Each line solver is covered: partition (K = 2, 8),
cyclic reduction, and Thomas per line and per system."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

RUN = r"""
import sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import emu_lib
emu_lib.use_emulator()
import numpy as np
from hydist_c_amd import abi, engine, scenarios as S
form, solver, param = {form}, {solver}, {param}
scens = []
for k in range(2):
    sc = S.meander_islands(72, 40, n_islands=3, seed=5 + k)
    sc.sediment_start = sc.bed_change_start = 4.0
    scens.append(sc)
b = engine.Batch(scens, form=form, sediment=True, bed_change=True, graph=True)
if solver is not None:
    for d in (0, 1):
        for hyd in (0, 1):
            engine.set_solver(b, d, hyd, solver, param)
b.init()
b.step(6)
out = {{f"{{k}}_{{f}}": b.get(k, f) for k in range(2) for f in ("U", "V", "Z", "FS", "DH", "KHOUOT")}}
np.savez({path!r}, **out)
b.close()
print("ok", flush=True)
"""

# (form, solver kind, param): the forms' defaults and every other solver
CASES = {
    "plain": ("abi.HD_FORM_FP64_PLAIN", "None", 0),
    "plain_seg": ("abi.HD_FORM_FP64_PLAIN", "engine.SOLVER_SEGMENTS", 0),
    "fast": ("abi.HD_FORM_FP64_FAST", "None", 0),
    "fast_part2": ("abi.HD_FORM_FP64_FAST", "engine.SOLVER_PARTITION", 2),
    "fast_part8": ("abi.HD_FORM_FP64_FAST", "engine.SOLVER_PARTITION", 8),
    "fast_cr": ("abi.HD_FORM_FP64_FAST", "engine.SOLVER_CR", 128),
    "fast_seg": ("abi.HD_FORM_FP64_FAST", "engine.SOLVER_SEGMENTS", 0),
    "fp32": ("abi.HD_FORM_FP32_REFINE", "None", 0),
    "fp32_cr": ("abi.HD_FORM_FP32_REFINE", "engine.SOLVER_CR", 256),
}


@pytest.fixture(scope="module")
def emu_built():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    build_emu.build()


def run(tmp_path, case, order):
    form, solver, param = CASES[case]
    path = str(tmp_path / f"{case}_{order.replace(':', '_')}.npz")
    code = RUN.format(root=ROOT, tests=HERE, form=form, solver=solver, param=param, path=path)
    env = dict(os.environ, HD_EMU_ORDER=order, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]
    with np.load(path) as z:
        return {k: z[k] for k in z.files}


def test_same_bits_in_any_order(emu_built, tmp_path):
    orders = ("forward", "reverse", "shuffle:7")
    jobs = [(c, o) for c in CASES for o in orders]
    with ThreadPoolExecutor(max_workers=4) as ex:  # four child processes at a time
        res = dict(zip(jobs, ex.map(lambda j: run(tmp_path, *j), jobs)))
    for case in CASES:
        a = res[(case, "forward")]
        for order in orders[1:]:
            b = res[(case, order)]
            assert a.keys() == b.keys()
            for k in a:
                assert np.array_equal(a[k].view(np.uint64) if a[k].dtype == np.float64 else a[k],
                                      b[k].view(np.uint64) if b[k].dtype == np.float64 else b[k]), (case, order, k)
