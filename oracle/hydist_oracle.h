/*
 * hydist_oracle.h - CPU FP64 oracle for the Hydist_C ADI timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU implementation is
 * compared against; it is never linked into or called by the product
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.
 *
 * The API mirrors include/hydist.h (same mesh/coeff/option structs, same
 * field ids) so a test can drive both with identical inputs.
 */
#ifndef HYDIST_ORACLE_H
#define HYDIST_ORACLE_H

#include "hydist.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ho_ctx ho_ctx;

/* Load_coeffs (loader.cpp:4-87), restated. */
int ho_load_coeffs(hd_coeffs *c);
/* Recompute Load_coeffs' derived members (dX2, Ks, Ufr, Dxr, wss, ...)
 * from the primary ones. */
int ho_finalize_coeffs(hd_coeffs *c);

int ho_set_threads(int n);
int ho_get_threads(void);

int ho_create(const hd_mesh *mesh, const hd_coeffs *coeffs, const hd_options *ops,
              const hd_config *cfg, ho_ctx **out);
int ho_init(ho_ctx *o);
int ho_update_boundary(ho_ctx *o, double t);
int ho_halfstep_uz(ho_ctx *o, double t);
int ho_halfstep_vz(ho_ctx *o, double t);
/* Stage blocks (engine.cpp:149-165, :258-282; sediment_transport.cpp
 * Calculate_Qb + BedLoad), as hd_sediment_j / hd_sediment_i / hd_bedload. */
int ho_sediment_j(ho_ctx *ctx, double t);
int ho_sediment_i(ho_ctx *ctx, double t);
int ho_bedload(ho_ctx *ctx, double t);
/* One kernel restatement on the current state, for closed-form checks:
 * "normalize_u", "normalize_v", "find_vth", "heso_k", "calculate_qb",
 * "bed_load", "uz_sweep" / "vz_sweep" (the implicit sweep of every line,
 * extraction included), "reset_states_vertical" and "find_limits" (the
 * segment tables from the current khouot, with the FS boundary
 * interpolation that runs beside it). */
int ho_run_stage(ho_ctx *ctx, const char *name);
int ho_step(ho_ctx *o, double *t);
int ho_run(ho_ctx *o, double *t, double Tmax);
int ho_copy_out(ho_ctx *o, int field, double *dst, size_t n);
int ho_copy_in(ho_ctx *o, int field, const double *src, size_t n);
int ho_copy_out_i32(ho_ctx *o, int field, int32_t *dst, size_t n);
int ho_copy_in_i32(ho_ctx *o, int field, const int32_t *src, size_t n);
int ho_copy_out_boundary(ho_ctx *o, int which, double *dst, size_t n);
int ho_copy_in_boundary(ho_ctx *o, int which, const double *src, size_t n);
int ho_get_hmax(ho_ctx *o, double out[4]);
int ho_set_hmax(ho_ctx *o, const double in[4]);
int ho_get_segments(ho_ctx *o, int dir, int32_t *counts, size_t ncounts, int32_t *first,
                    int32_t *last, size_t cap, int64_t *total);
int ho_mass_balance(ho_ctx *o, double *volume, double *inflow);
int ho_destroy(ho_ctx *o);
const char *ho_last_error(ho_ctx *o);

/* Stage-level pieces, exposed for unit tests. */

/* tridiag (UVZSolver_multithread.cpp:19-43): Thomas without pivoting on
 * rows 0..sn. Ap/Bp are caller scratch of sn+1 entries. */
int ho_thomas(int sn, const double *AA, const double *BB, const double *CC, const double *DD,
              double *x, double *Ap, double *Bp);

/* Run-length wet-segment rule of Find_Calculation_limits_Horizontal
 * (support_funcs.cpp:260-288) on one line: wet[k] for k in [0, end+1]
 * holds khouot values along the line, end is M (or N).  Writes up to cap
 * segments and returns the count. */
int ho_segment_line(const int32_t *line, int end, int32_t *first, int32_t *last, int cap);

/* Assemble the interleaved (z,u) system of one hydraulic segment and
 * solve it (stage A4..A7 on a single line) - exposed so tests can
 * check the assembled system against its solution. dir 1 = UZ (column
 * `line`), dir 0 = VZ (row `line`).  Writes sn+1 rows into AA..DD and x
 * (capacity cap) and returns sn (or a negative status). */
int ho_line_system(ho_ctx *o, int dir, int line, int seg, double *AA, double *BB, double *CC,
                   double *DD, double *x, int cap);

/* Per-position coefficients of one hydraulic segment, as the preindex
 * kernels form them (_vzSolver_calculate_preindex,
 * UVZSolver_multithread.cpp:142-208, dir 0; _uzSolver_calculate_preindex
 * :210-278, dir 1): out[a * cap + p] for a = f1, f2, f3, f5, a2, c2, d2 and
 * p the position along the line (cap >= line length + 4; positions outside
 * the segment are 0).  Returns the segment's last position. */
int ho_line_coeffs(ho_ctx *o, int dir, int line, int seg, double *out, int cap);

/* The FS rows of one segment (dir 1: _FSj_calculate__mactrix_coeff,
 * sediment_transport.cpp:204-313; dir 0: _FSi_, :51-164) at time t, row k
 * for position first + 1 + k; nothing is solved.  Returns sn (-1 for a
 * segment of fewer than 3 cells). */
int ho_fs_line_system(ho_ctx *o, int dir, int line, int seg, double t, double s_start, double *AA,
                      double *BB, double *CC, double *DD, int cap);

/* source_chs / source_nchs (sediment_transport.cpp:11-49). */
double ho_source_chs(const hd_coeffs *c, double wsm, double Zf, double fs, double fw, double vth,
                     double h_moi);
double ho_source_nchs(const hd_coeffs *c, double wsm, double Zf, double Uf, double h_moi, double fs);

#ifdef __cplusplus
}
#endif

#endif
