"""bench.py's bookkeeping on the CPU: stage names map to the kernel symbols
rocprofv3 reports (so the roofline line names the profiled kernel), and
the algorithmic byte counts follow the stated per-position figures."""
import types

import bench


SC = types.SimpleNamespace(M=2048, N=1024)


def test_solver_stage_symbols_follow_the_tuned_choice():
    chosen = ["part/4", "part/8", "cr/256", "cr/0"]
    assert bench.kernel_symbol("k_part<f64,uz>", SC, chosen, 1) == "k_part<double, false, true, 4, false, false>"
    assert bench.kernel_symbol("k_part<f64,vz>", SC, chosen, 1) == "k_part<double, false, true, 8, true, true>"
    assert bench.kernel_symbol("k_part<f32r,vz>", SC, chosen, 1) == "k_part<float, true, true, 8, true, true>"
    # row-line solves extract their solution themselves (EXTRACT = true)
    assert bench.kernel_symbol("k_part<f64,fs_i>", SC, ["part/4", "part/8", "part/8", "part/2"], 1) == \
        "k_part<double, false, false, 2, true, true>"
    # cyclic reduction: 1023 unknowns over 256 threads -> 4 per thread
    assert bench.kernel_symbol("k_cr_reg<f64,fs_j>", SC, chosen, 1) == "k_cr_reg<double, false, false, 4>"
    # default workgroup rule (512 above 1024 unknowns): 2047 / 512 -> 4
    assert bench.kernel_symbol("k_cr_reg<f64,fs_i>", SC, chosen, 1) == "k_cr_reg<double, false, false, 4>"
    # without a tuned choice: the default K (4 up to 2048 unknowns)
    assert bench.kernel_symbol("k_part<f64,vz>", SC, None, 1) == "k_part<double, false, true, 4, true, true>"


def test_assembly_and_extraction_symbols_by_layout():
    assert bench.kernel_symbol("k_hyd_assemble<uz>", SC, None, 1) == "k_hyd_assemble_cols<1>"
    assert bench.kernel_symbol("k_hyd_assemble<vz>", SC, None, 1) == "k_hyd_assemble<0, 1>"
    assert bench.kernel_symbol("k_hyd_assemble<uz>", SC, None, 0) == "k_hyd_assemble<1, 0>"
    assert bench.kernel_symbol("k_fs_assemble<j>", SC, None, 1) == "k_fs_assemble_cols<1>"
    assert bench.kernel_symbol("k_fs_extract<i>", SC, None, 0) == "k_fs_extract<0, 0>"
    assert bench.kernel_symbol("k_hyd_extract<uz>", SC, None, 1) == "k_hyd_extract_cols"
    assert bench.kernel_symbol("k_thomas<vz>", SC, None, 0) == "k_thomas"
    assert bench.kernel_symbol("k_thomas_seg<fs_j>", SC, None, 0) == "k_thomas_seg"
    assert bench.kernel_symbol("k_post_a", SC, None, 1) == "k_post_a"
    # the post passes that also form VTH, Kx, Ky run the same kernels
    assert bench.kernel_symbol("k_post_a+vth", SC, None, 1) == "k_post_a"
    assert bench.kernel_symbol("k_post_b+vth", SC, None, 1) == "k_post_b"


def test_stage_bytes_all_positions():
    # every position of every line, identity rows included: hydraulic solve
    # 80 B per position, (M-1)(N-1) positions
    pos = (SC.M - 1) * (SC.N - 1)
    u = {"pos_0": pos, "pos_1": pos}
    assert bench.stage_bytes("k_part<f64,uz>", u, SC, all_positions=True) == 80.0 * pos
    assert bench.stage_bytes("k_cr_reg<f32r,vz>", u, SC, all_positions=True) == 80.0 * pos
    assert bench.stage_bytes("k_part<f64,vz>", u, SC, all_positions=True) == 112.0 * pos
    assert bench.stage_bytes("k_part<f64,fs_j>", u, SC, all_positions=True) == 40.0 * pos
    assert bench.stage_bytes("k_part<f64,fs_i>", u, SC, all_positions=True) == 104.0 * pos
    assert bench.stage_bytes("k_hyd_assemble<uz>", u, SC, all_positions=True) == 160.0 * pos
    cells = (SC.M + 3) * (SC.N + 3)
    assert bench.stage_bytes("k_segment", u, SC) == 8.0 * cells
    # with VTH, Kx, Ky: three fields out and H_moi (and htaiz) in on top
    assert bench.stage_bytes("k_post_a+vth", u, SC) == bench.stage_bytes("k_post_a", u, SC) + 32.0 * cells
    assert bench.stage_bytes("k_post_b+vth", u, SC) == bench.stage_bytes("k_post_b", u, SC) + 40.0 * cells
    assert bench.stage_bytes("no_such_kernel", u, SC) is None


class FakeSim:
    """Segment tables of two row segments (L = 5, 2) and one column
    segment (L = 4)."""
    import numpy as np

    def segments(self, d):
        np = self.np
        if d == 0:
            return np.array([0, 2]), np.array([2, 10]), np.array([6, 11])
        return np.array([1]), np.array([3]), np.array([6])


def test_line_units_and_wet_pricing():
    u = bench.line_units(FakeSim(), SC)
    assert u["wet_0"] == 7 and u["il_0"] == 9 + 3 and u["fs_0"] == 3 and u["fsc_0"] == 5
    assert u["wet_1"] == 4 and u["il_1"] == 7 and u["fs_1"] == 2 and u["fsc_1"] == 4
    # solves: 40 B per interleaved unknown of the wet segments
    assert bench.stage_bytes("k_part<f64,uz>", u, SC) == 40.0 * 7
    assert bench.stage_bytes("k_thomas<vz>", u, SC) == 40.0 * 12
    assert bench.stage_bytes("k_part<f64,fs_j>", u, SC) == 40.0 * 2
    # one-kernel sweeps: the state of the wet positions in, results out
    assert bench.stage_bytes("k_part<f64,vz>", u, SC) == 112.0 * 7
    assert bench.stage_bytes("k_part<f64,fs_i>", u, SC) == 96.0 * 3 + 8.0 * 5
    assert bench.stage_bytes("k_hyd_assemble<uz>", u, SC) == 160.0 * 4
    assert bench.stage_bytes("k_hyd_extract<uz>", u, SC) == 32.0 * 4
