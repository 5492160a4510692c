// hd_kernels.h - declarations of the gfx950 kernels (hd_kernels.hip).
#pragma once

#include "hd_device.h"

namespace hd {

__global__ void k_onetime_init(const DevState *S);
// k_segment: workgroups of SEG_COLS x SEG_GROUPS threads; a row block takes
// SEG_ROWS_PER_BLOCK rows (one per wavefront), a column block SEG_COLS
// columns
constexpr int SEG_COLS = 16, SEG_GROUPS = 64;
constexpr int SEG_ROWS_PER_BLOCK = 4;
__global__ void k_segment(const DevState *S, int row_blocks, int col_blocks);
__global__ void k_htuongdoi(const DevState *S);
__global__ void k_preprocess(const DevState *S);
__global__ void k_boundary(const DevState *S, double t, int from_clock);
__global__ void k_clock(double *clock, double v, int advance);
template <bool ISU> __global__ void k_smooth(const DevState *S);
__global__ void k_solve_v(const DevState *S);
__global__ void k_solve_u(const DevState *S);
template <int DIR> __global__ void k_margin(const DevState *S);
__global__ void k_post_a(const DevState *S, int vth);
__global__ void k_pre_reset(const DevState *S);
__global__ void k_rewet(const DevState *S);
__global__ void k_dry(const DevState *S);
__global__ void k_post_b(const DevState *S, int vth);
__global__ void k_interp_fs(const DevState *S);
__global__ void k_vth_k(const DevState *S);
__global__ void k_update_fs(const DevState *S);
__global__ void k_qb(const DevState *S);
template <bool UPDATE_FS> __global__ void k_bedload(const DevState *S);
__global__ void k_volume(const DevState *S);
__global__ void k_watchdog(const DevState *S, int *flag);
__global__ void k_set_flag(int *flag, int v);

}  // namespace hd
