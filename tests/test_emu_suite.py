"""Small cases of the GPU tests, run on the CPU: the library's kernels
built against the emulated HIP runtime (tests/emu: every lane on the host,
device buffers and dynamic LDS against guard pages) and checked by the
same test code as on the GPU (HD_EMU=1), in a child process so the
emulated library never meets the GPU build in one interpreter.  This puts
the kernels' indexing and arithmetic (bitwise plain form, fast forms at
their bars, every line solver, the tridiagonal entry points, degenerate
meshes, the serving loop with its worker and context-making threads)
under the CPU suite; timing and occupancy stay GPU questions."""
import os
import re
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

FILES = ["test_gpu_gtsv.py", "test_zz_no_alloc.py", "test_gpu_parity.py", "test_gpu_fast.py",
         "test_gpu_edge_meshes.py", "test_gpu_server.py", "test_gpu_determinism.py",
         "test_gpu_random_parity.py", "test_gpu_thomas_seg.py", "test_gpu_api_contract.py",
         "test_gpu_restart.py"]
# the cases that finish in about a minute on the emulator together (four
# processes)
SELECT = " or ".join([
    "gtsv", "noalloc", "out_of_order_and_bad", "bad_ids_sizes", "create_refuses", "restart_continues", "restore_rejects", "test_server", "two_runs", "footprint_is", "(random_scenario and ([0- or [1- or [2-))",
    "(thomas_per_system and ([box] or [onecell] or [flats]))", "plain_tuning", "fp64_forms", "in_the_fast_form", "fast_form_thomas",
    "(test_plain_fp64_bitwise and (12_0 or 12_1 or -20] or -25]))",
    "half_steps", "sediment_and_bed_change", "lake_at_rest", "drying_threshold", "stage_entry_points", "q11_literal",
    "init_matches", "mass_balance_gpu", "output_callback",
    "(test_fast_form_close_to_oracle and 10_)", "test_fast_form_sediment",
    "(test_every_line_solver_close_to_oracle and part4)",
    "edge_mesh_plain", "(edge_mesh_fast and (stripes or onecell or alland))",
])


# a cross-section of the kernels for the reversed-order run
REVERSE = " or ".join([
    "gtsv", "noalloc", "(random_scenario and ([0- or [1-))", "(thomas_per_system and [box])",
    "(test_plain_fp64_bitwise and 12_0)", "sediment_and_bed_change", "(test_fast_form_close_to_oracle and 10_)",
    "test_fast_form_sediment", "(test_every_line_solver_close_to_oracle and part4)",
    "(edge_mesh_fast and (stripes or onecell))",
])


@pytest.fixture(scope="module")
def emu_built():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    build_emu.build()


def run_cases(select, minimum, **env_extra):
    env = dict(os.environ, HD_EMU="1", OMP_NUM_THREADS="2", **env_extra)  # the oracle's threads, per worker
    # four worker processes (pytest-xdist): the cases are independent
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-n", "4", "-p", "no:cacheprovider", "-k", select] + \
          [os.path.join(HERE, f) for f in FILES]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=os.path.dirname(HERE), timeout=900)
    tail = r.stdout[-3000:] + r.stderr[-2000:]
    assert r.returncode == 0, tail
    m = re.search(r"(\d+) passed", r.stdout)
    assert m and int(m.group(1)) >= minimum, tail  # the selection, not an empty one
    print(r.stdout.strip().splitlines()[-1])


def test_gpu_suite_small_cases_on_the_emulator(emu_built):
    run_cases(SELECT, 90)


def test_results_do_not_depend_on_lane_or_workgroup_order(emu_built):
    """The same checks with every launch's workgroups, and the lanes of each
    pass between barriers, run in reverse order (HD_EMU_ORDER=reverse): a
    kernel that reads LDS another lane writes without a barrier between, or
    global memory another workgroup of the same launch writes, gives other
    results and fails them (the plain form against the oracle bit for
    bit)."""
    run_cases(REVERSE, 20, HD_EMU_ORDER="reverse")
