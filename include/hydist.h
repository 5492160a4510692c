/*
 * hydist.h - C ABI of the MI355X (gfx950) implementation of the Hydist_C
 * ADI shallow-water + suspended-sediment timestep.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * an int status (0 = OK, negative = HD_E_*) and never throws.  A context
 * (hd_ctx) owns every device buffer of one scenario and one HIP stream;
 * calls on a context are asynchronous until hd_sync() unless stated.
 *
 * Field layout (the reference's, loader.cpp:400, support_funcs.cpp:39-43):
 * 2-D fields are row-major [i][j], i in [0, N+2] (rows), j in [0, M+2]
 * (columns, unit stride).  Compute cells are i in [2, N], j in [2, M].
 * Host buffers passed to the API are dense (N+3) x (M+3); the device copy
 * uses a padded pitch chosen by the library.
 *
 * Each declaration names the reference interface it replaces (file:line,
 * in the Hydist_C sources).
 */
#ifndef HYDIST_H
#define HYDIST_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
#define HD_OK 0
#define HD_E_INVALID (-1)   /* bad argument / shape */
#define HD_E_NOMEM (-2)     /* device allocation failed / over budget */
#define HD_E_HIP (-3)       /* HIP runtime error (message in hd_last_error) */
#define HD_E_NUMERIC (-4)   /* NaN/Inf watchdog tripped */
#define HD_E_STATE (-5)     /* call out of order (e.g. step before init) */
#define HD_E_NODEV (-6)     /* no usable gfx950 device */
#define HD_E_FAULT (-7)     /* injected fault (test hook) */
#define HD_E_CANCELLED (-8) /* serving request cancelled */
#define HD_E_UNSUPPORTED (-9)

/* ------------------------------------------------------------------ */
/* Constant_Coeffs: byte-for-byte the reference struct (engine.h:39-52) */
/* ------------------------------------------------------------------ */
typedef struct hd_coeffs {
    double dX, dY, dT, dXbp, dYbp, dX2, dY2, dTchia2dX, dTchia2dY, HaiChiadT;
    double QuyDoiTime, QuyDoiPi;
    double NANGDAY, H_TINH, heso, mu_mn;
    double kenhhepng, kenhhepd;
    double Wind, huonggio, Zbandau;
    double NDnen, NDbphai, NDbtrai, NDbtren, NDbduoi;
    double Tod, Toe, hstoe, ghtoe, Mbochat;
    double ro, ros, dm, muy, d90, Dorong, Sx, sohn, soha, sohn1;
    double CORIOLIS_FORCE, g, Ks;
    double Windx, Windy, Ufr, Dxr, wss;
} hd_coeffs;

/* ------------------------------------------------------------------ */
/* Options: the data members of the reference struct (engine.h:16-37)  */
/* in the same order and types (C bool == C++ bool on x86-64).         */
/* ------------------------------------------------------------------ */
typedef struct hd_options {
    double Tmax, t_start, sediment_start, bed_change_start;
    int interval;
    bool kenhhepd, kenhhepng;
    int M, N;
    bool ketdinh, channel, debug, plot;
    int total_time;
} hd_options;

/* Boundary sides, in the order Update_Boundary_Value visits them
 * (support_funcs.cpp:692-725).  boundary_type[side] == 2 means a discharge
 * (Q) boundary, anything else a water-level (Z) boundary
 * (loader.cpp:158-164). */
enum { HD_SIDE_UP = 0, HD_SIDE_DOWN = 1, HD_SIDE_LEFT = 2, HD_SIDE_RIGHT = 3 };

/* ------------------------------------------------------------------ */
/* Mesh + inputs of one scenario (replaces load_inputs' 13 vectors,    */
/* loader.cpp:120-168, and Host_arrays, engine.h:80-88).               */
/* ------------------------------------------------------------------ */
typedef struct hd_mesh {
    int M, N;                        /* arrays are (N+3) x (M+3), row-major */
    const double *h;                 /* bandodosau: depth at cell corners (required) */
    const double *hsnham;            /* Manning-like roughness (required) */
    const double *VISCOIDX;          /* horizontal eddy viscosity A (required) */
    const double *Fw;                /* bed friction factor (required when sediment is on; NULL = 0) */
    const double *Tsxw, *Tsyw;       /* wind stress fields (NULL = 0) */
    const double *u0, *v0, *z0, *FS0; /* initial state (NULL = 0) */
    int total_time;                  /* samples (hourly) per boundary series, >= hours run + 2 */
    int boundary_type[4];            /* per side, see HD_SIDE_* */
    const double *bc[4];             /* per side: total_time hourly samples (Z in m or Q in m^3/s) */
    const double *CC[4];             /* per side: total_time hourly FS samples (kg/m^3); NULL = 0 */
} hd_mesh;

/* Numerical forms of the implicit sweeps. */
enum {
    HD_FORM_FP64_PLAIN = 0,  /* interleaved (z,u) Thomas in the reference's order: bitwise FP64 */
    HD_FORM_FP64_FAST = 1,   /* z-only (Schur) system, partitioned parallel solve, FP64 */
    HD_FORM_FP32_REFINE = 2  /* z-only system solved in FP32 + one FP64 residual refinement */
};

/* Semantics switches for reference behaviours that are defined differently
 * by default (each documented where it is implemented). */
#define HD_QUIRK_Q11_LITERAL 0x1 /* open-start segments test the start side's Q flag at their end,
                                    exactly as UVZSolver_multithread.cpp:438 */

typedef struct hd_config {
    int form;          /* HD_FORM_* */
    int sediment;      /* run the FS sweeps once t >= sediment_start (engine.cpp:149,258) */
    int bed_change;    /* accumulate dH (Calculate_Qb + BedLoad) once t >= bed_change_start */
    int quirks;        /* HD_QUIRK_* bits */
    int device;        /* HIP device ordinal */
    int dense_layout;  /* 1: device fields use the reference's width M+3 (no row padding) */
} hd_config;

/* Field ids for hd_copy_out / hd_copy_in. */
enum {
    HD_F_U = 0, HD_F_V, HD_F_Z, HD_F_TU, HD_F_TV, HD_F_TZ,
    HD_F_HMOI, HD_F_HTDU, HD_F_HTDV, HD_F_FS, HD_F_TFS,
    HD_F_VTH, HD_F_KX, HD_F_KY, HD_F_QBX, HD_F_QBY, HD_F_DH,
    HD_F_HTAIZ, HD_F_HTAIZ_BD, HD_F_KX1, HD_F_KY1,
    HD_F_H, HD_F_HSNHAM, HD_F_VISCOIDX, HD_F_FW,
    HD_F_KHOUOT, /* int32 field: copy with hd_copy_out_i32 / hd_copy_in_i32 */
    HD_F_COUNT
};

/* Boundary work arrays (device; Argument_Pointers ubt/ubp/vbt/vbd/hi,
 * engine.h:61-65, sizes from loader.cpp:306-320). */
enum { HD_B_UBT = 0, HD_B_UBP, HD_B_VBT, HD_B_VBD, HD_B_HI, HD_B_COUNT };

typedef struct hd_ctx hd_ctx;

/* Load_coeffs (loader.cpp:4-87): fills the reference's constants. */
int hd_load_coeffs(hd_coeffs *c);

/* Recomputes the members Load_coeffs derives (dXbp, dX2, dTchia2dX,
 * HaiChiadT, Ks, Windx/Windy, Ufr, Dxr, wss; loader.cpp:9-17,79-86) from
 * the primary ones, for callers that change dX, dT, dm, ... */
int hd_finalize_coeffs(hd_coeffs *c);

/* attribute_arrays_memory_alloc + supporting_arrays_alloc
 * (loader.cpp:224-375, main.cpp:216-233): allocate every device buffer of
 * one scenario and upload its inputs.  HD_E_INVALID for M or N below 4, a
 * missing required array or boundary series, fewer than 2 series samples,
 * an unknown form or quirk bit, or dX, dY, dT not positive and finite;
 * ops->M and ops->N are taken from the mesh. */
int hd_create(const hd_mesh *mesh, const hd_coeffs *coeffs, const hd_options *ops,
              const hd_config *cfg, hd_ctx **out);

/* main.cpp:257-299: Onetime_init, Find_Calculation_limits_{Horizontal,
 * Vertical}, Htuongdoi and preprocess_data, on the device. */
int hd_init(hd_ctx *ctx);

/* update_boundary_at_t (engine.cpp:22-36) -> Update_Boundary_Value.
 * HD_E_STATE before hd_init, which derives the Q sides' weights. */
int hd_update_boundary(hd_ctx *ctx, double t);

/* First half-step (engine.cpp:66-170): t is the time after t += dT/2.
 * Runs the boundary update, the UZ implicit sweep, explicit V, smoothing,
 * wet-depth update, segmentation, Htuongdoi and, when enabled and
 * t >= sediment_start, the FS sweep along i. */
int hd_halfstep_uz(hd_ctx *ctx, double t);

/* Second half-step (engine.cpp:176-288), including Reset_states_vertical,
 * the FS sweep along j and (when enabled) the bed change. */
int hd_halfstep_vz(hd_ctx *ctx, double t);

/* One full dT from *t (both half-steps); advances *t by dT. */
int hd_step(hd_ctx *ctx, double *t);

/* Hydraulic_Calculation's loop (engine.cpp:61-295): steps while t < Tmax
 * (Tmax truncated to int as engine.cpp:62).  *t holds the start time on
 * entry and the final time on return. */
int hd_run(hd_ctx *ctx, double *t, double Tmax);

/* Hydraulic_Calculation's loop with output: as hd_run, and after every
 * step at which t is a whole multiple of `interval` seconds
 * (engine.cpp:145,243) the stream is synchronised and cb(ctx, t, user) is
 * called (it may copy fields out); a nonzero return stops the run with
 * HD_E_CANCELLED. */
typedef int (*hd_output_cb)(hd_ctx *ctx, double t, void *user);
int hd_run_cb(hd_ctx *ctx, double *t, double Tmax, int interval, hd_output_cb cb, void *user);

/* Stage entry points for per-stage use and parity.  hd_sediment_j: the
 * first half-step's FS block, Find_VTH + hesoK, Scan_FSj, the solve,
 * FSj_extract_solution and Update_FS (engine.cpp:149-165);
 * hd_sediment_i: the second half-step's (engine.cpp:258-282);
 * hd_bedload: Calculate_Qb + BedLoad (sediment_transport.cpp:352-537,
 * intended form).  t is the half-step's time. */
int hd_sediment_j(hd_ctx *ctx, double t);
int hd_sediment_i(hd_ctx *ctx, double t);
int hd_bedload(hd_ctx *ctx, double t);

/* Block until all work queued on the context has finished. */
int hd_sync(hd_ctx *ctx);

/* save_result / save_FS (loader.cpp:397-434) data access: dense
 * (N+3) x (M+3) copies between host and the device field. n is the
 * element count of the host buffer. */
int hd_copy_out(hd_ctx *ctx, int field, double *dst, size_t n);
int hd_copy_in(hd_ctx *ctx, int field, const double *src, size_t n);
int hd_copy_out_i32(hd_ctx *ctx, int field, int32_t *dst, size_t n);
int hd_copy_in_i32(hd_ctx *ctx, int field, const int32_t *src, size_t n);

/* Boundary work arrays: HD_B_UBT/UBP have M+3 entries, VBT/VBD N+3,
 * HD_B_HI 2*(M+N+6) (layout of preprocess_data, support_funcs.cpp:505-570). */
int hd_copy_out_boundary(hd_ctx *ctx, int which, double *dst, size_t n);
int hd_copy_in_boundary(hd_ctx *ctx, int which, const double *src, size_t n);
/* hmax_u, hmax_d, hmax_l, hmax_r (Argument_Pointers, engine.h:57). */
int hd_get_hmax(hd_ctx *ctx, double out[4]);
int hd_set_hmax(hd_ctx *ctx, const double in[4]);

/* Wet-segment tables (daui/cuoii/moci and dauj/cuoij/mocj,
 * support_funcs.cpp:238-343) as CSR: dir 0 = along j per row i (lines
 * i = 0..N+2), dir 1 = along i per column j (lines j = 0..M+2).
 * counts[line] receives the number of segments, first/last the segment
 * bounds in line order; cap is the capacity of first/last.  *total
 * receives the total segment count (may exceed cap: then only cap are
 * written). */
int hd_get_segments(hd_ctx *ctx, int dir, int32_t *counts, size_t ncounts,
                    int32_t *first, int32_t *last, size_t cap, int64_t *total);

/* Domain water volume sum over wet compute cells of (htaiz + z)*dX*dY and
 * the running sum of boundary inflow volume, both accumulated in FP64 on
 * the device in a fixed order (either output may be NULL).  HD_E_STATE
 * before hd_init. */
int hd_mass_balance(hd_ctx *ctx, double *volume, double *inflow);

/* NaN/Inf watchdog (the reference checks only for launch errors after
 * every kernel, synch_and_check, engine.cpp:13-19).  With every > 0, after
 * every `every`-th full step (hd_step, hd_run, hd_run_cb; eager or graph)
 * one kernel checks z, u and v of every wet compute cell of every
 * scenario.  hd_run and hd_run_cb wait for each check and stop with
 * HD_E_NUMERIC at the first that finds a non-finite value; after hd_step
 * the finding is returned by the next hd_sync.  Every later hd_sync,
 * hd_run and hd_run_cb returns HD_E_NUMERIC too, until the watchdog is
 * set again (which clears the finding).  every = 0 switches it off (the
 * default). */
int hd_set_watchdog(hd_ctx *ctx, int every);

/* Fault injection for failure-handling tests (no counterpart in the
 * reference): stage 1 makes the next UZ half-step, stage 2 the next VZ
 * half-step, return HD_E_FAULT before it launches anything; 0 clears it.
 * The fault stays armed until cleared. */
int hd_inject_fault(hd_ctx *ctx, int stage);

/* Frees every buffer of the context. */
int hd_destroy(hd_ctx *ctx);

/* Last error message of the context (or of the library when ctx is NULL). */
const char *hd_last_error(hd_ctx *ctx);

/* Library/device information: writes a short text (arch, CU count,
 * build flags) into buf. */
int hd_device_info(char *buf, size_t n);

/* Stage timing (no counterpart in the reference, which has no timers):
 * when on, every kernel launch of the context is bracketed by HIP events
 * on its stream.  hd_stage_stats writes {"stage": [launches, total_ms], ...}
 * as JSON into buf (optionally clearing the counters). */
int hd_set_profiling(hd_ctx *ctx, int on);
int hd_stage_stats(hd_ctx *ctx, char *buf, size_t n, int reset);

/* ------------------------------------------------------------------ */
/* Ensembles.  The reference runs one scenario per process (main.cpp); */
/* these calls run n independent scenarios of one mesh shape in one    */
/* context: every stage is one launch over all of them (the scenario   */
/* is the grid's outer dimension), and hd_step / hd_run / hd_init /    */
/* hd_halfstep_* advance all of them together.  Scenarios may differ   */
/* in every field and series; M, N and total_time are shared, and so   */
/* are coefficients, options and form.  Scenario 0 is the one the      */
/* single-scenario calls (hd_copy_out, hd_get_segments, ...) address.  */
/* ------------------------------------------------------------------ */
int hd_batch_create(const hd_mesh *meshes, int n, const hd_coeffs *coeffs, const hd_options *ops,
                    const hd_config *cfg, hd_ctx **out);
int hd_batch_size(hd_ctx *ctx, int *n);
/* As hd_copy_out / hd_copy_in for scenario `scen`; HD_F_KHOUOT is
 * returned as doubles. */
int hd_batch_copy_out(hd_ctx *ctx, int scen, int field, double *dst, size_t n);
int hd_batch_copy_in(hd_ctx *ctx, int scen, int field, const double *src, size_t n);

/* Whole-step graphs (engine.cpp:61-295 issues each stage as a separate
 * launch followed by a device sync).  When on, hd_step / hd_run replay one
 * captured hipGraph per step: one graph per combination of optional
 * stages (sediment in either half-step, bed change), captured the first
 * time it is needed.  Results are identical to eager launches. */
int hd_set_graph(hd_ctx *ctx, int on);

/* Line solvers of the fast forms (they replace the Thomas solve,
 * UVZSolver_multithread.cpp:19-43, and the unused SPIKE solver,
 * gtsv.cpp:88-193).  One workgroup solves one line:
 *   HD_SOLVER_PARTITION: each thread eliminates K consecutive unknowns in
 *     registers, the one interface unknown per thread is solved by
 *     parallel cyclic reduction in LDS (param = K: 2, 4 or 8; lines up to
 *     512*K unknowns);
 *   HD_SOLVER_CR: cyclic reduction of the whole line in LDS (param = the
 *     workgroup size, a multiple of 64 up to 512; lines up to 8 x that).
 * HD_SOLVER_DEFAULT is the partition solver with K = 4 (K = 8 above 2048
 * unknowns).  The two solvers, and the partition solver at different K,
 * round differently (agreement ~1e-15 relative in FP64); the CR workgroup
 * size does not change any bit. */
/* The plain form always solves with Thomas in the reference's order
 * (tridiag, UVZSolver_multithread.cpp:19-43): by default one lane walks a
 * whole line, its systems one after the other; HD_SOLVER_SEGMENTS
 * (param 0) gives every system of the sweep a lane of its own.  Both give
 * the same bits.  The fast FP64 form takes HD_SOLVER_SEGMENTS too: the
 * same FP64 Thomas on the fast form's assembled systems, one more choice
 * rounding differently from its other solvers (FP32 + refine: rejected). */
enum { HD_SOLVER_DEFAULT = 0, HD_SOLVER_PARTITION = 1, HD_SOLVER_CR = 2, HD_SOLVER_SEGMENTS = 3 };

/* Autotuning (no counterpart in the reference).  Times each line solver
 * candidate (fast forms: partition at K = 2, 4, 8, cyclic reduction at
 * 128, 256, 512 threads, and in the fast FP64 form Thomas per system;
 * plain form: Thomas per line or per system) on
 * every sweep (UZ, VZ, and the FS sweeps when sediment is on) from the
 * context's current state, keeps the fastest per sweep, and writes a JSON
 * report (ms per candidate and the choices) into report.  In the fast
 * forms the choice can change results at rounding level (see above); it
 * is deterministic once made.  In the plain form it changes no bit.  hd_set_solver sets one choice by hand
 * (dir 1 = UZ, 0 = VZ; hydraulic 1 = the z/u sweep, 0 = the FS sweep);
 * hd_set_solver_threads(ctx, dir, hyd, t) is hd_set_solver(ctx, dir, hyd,
 * HD_SOLVER_CR, t). */
int hd_autotune(hd_ctx *ctx, char *report, size_t n);
int hd_set_solver(hd_ctx *ctx, int dir, int hydraulic, int kind, int param);
int hd_set_solver_threads(hd_ctx *ctx, int dir, int hydraulic, int threads);

/* Batched tridiagonal solve on device buffers, the drop-in for
 * gtsv_spike_partial_diag_pivot_v1 (gtsv.cpp:88, declared gtsv.h:7; same
 * argument meaning: dl/d/du the sub-, main and super-diagonals, b the
 * right-hand side, x the solution, m rows; dl[0] and du[m-1] are ignored)
 * extended to nsys systems stored ld elements apart.  Solved by the
 * partition solver without pivoting, so the systems must be diagonally
 * dominant (every system this library forms is); precision 0 = FP64,
 * 1 = FP32 with one FP64 refinement.  Synchronous on stream (a
 * hipStream_t, NULL = the null stream).  m <= 4096. */
int hd_gtsv_batch(const double *dl, const double *d, const double *du, const double *b, double *x,
                  int m, int nsys, int ld, int precision, void *stream);

/* One tridiagonal system of any length on device buffers: the drop-in for
 * gtsv_spike_partial_diag_pivot_v1 (gtsv.cpp:88-193), which splits one
 * system of m rows over many workgroups and solves their interface system
 * in one block, allocating its work buffers on every call (:111-147).
 * Same argument meaning as hd_gtsv_batch for nsys = 1 (dl[0], du[m-1]
 * ignored), 1 <= m <= 4096 * 4096.  Above 4096 rows the system is cut
 * into chunks of 4096 rows, each chunk's interior is solved by the
 * partition solver for the right-hand side and two spikes, and the
 * interface unknowns (one per chunk) by one more partition solve.  Work
 * memory (4 m + 5 m / 4096 doubles) is allocated and freed on `stream`
 * for each call.  Synchronises the stream before returning.  Diagonally
 * dominant systems, as hd_gtsv_batch. */
int hd_gtsv_long(const double *dl, const double *d, const double *du, const double *b, double *x,
                 int m, int precision, void *stream);

/* hd_gtsv_long without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops that must not call hipMalloc, harnesses
 * that forbid it).  Same arguments and results (bit for bit) as
 * hd_gtsv_long, for m <= HD_GTSV_NOALLOC_MAX_M; larger m is rejected with
 * HD_E_INVALID.  Its work memory is a fixed device array of
 * 4 HD_GTSV_NOALLOC_MAX_M + 320 doubles (8.4 MB) in the library, one per
 * device, held until the process ends and shared by every call on that
 * device: calls that overlap on one device (other streams, other host
 * threads) race on it and must be serialised by the caller. */
#define HD_GTSV_NOALLOC_MAX_M 262144
int hd_gtsv_long_noalloc(const double *dl, const double *d, const double *du, const double *b, double *x,
                         int m, int precision, void *stream);

/* Checkpoint / restart.  Replaces the reference's stubs (the -c/-T
 * options, main.cpp:283-284, loader.cpp:172-185): the snapshot holds every
 * buffer whose contents carry from one step to the next (state, derived
 * fields, boundary arrays, segment tables, inflow accumulator) and the
 * time t, in the device layout; it restores into a context created from
 * the same mesh and config, and the run then continues bit for bit as if
 * never interrupted.  The header keeps a checksum of itself and the payload:
 * hd_restore refuses (HD_E_INVALID) a snapshot of another mesh or layout
 * and a damaged or truncated one, leaving the context as it was. */
int hd_snapshot_size(hd_ctx *ctx, size_t *bytes);
int hd_snapshot(hd_ctx *ctx, void *buf, size_t n, double t);
int hd_restore(hd_ctx *ctx, const void *buf, size_t n, double *t);
int hd_save_checkpoint(hd_ctx *ctx, const char *path, double t);
int hd_load_checkpoint(hd_ctx *ctx, const char *path, double *t);

/* Device bytes a context of nscen scenarios of this mesh allocates
 * (hd_create / hd_batch_create), and what a context holds.  The
 * reference allocates without accounting (loader.cpp:224-375). */
int hd_footprint(const hd_mesh *mesh, const hd_config *cfg, int nscen, size_t *bytes);
int hd_memory_bytes(hd_ctx *ctx, size_t *bytes);

/* ------------------------------------------------------------------ */
/* Serving loop.  No counterpart in the reference (one scenario per    */
/* process, main.cpp): what-if runs are submitted as they arrive; a    */
/* worker thread admits them in order while their footprint fits the   */
/* HBM budget and a slot is free, keeps each admitted run resident in  */
/* a context and stream (made on a helper thread), and advances all    */
/* admitted runs together in rounds: the largest takes                 */
/* steps_per_quantum steps per round, a smaller one up to 4x as many,  */
/* in inverse proportion to its cell count (their kernels overlap on   */
/* the device).  Queued runs of one mesh size, duration, form, options */
/* and coefficients are admitted together into one batched context     */
/* (up to 16), so each launch serves all of them.  A run's result does */
/* not depend on what else runs.                                       */
/* ------------------------------------------------------------------ */
typedef struct hd_server hd_server;

typedef struct hd_server_config {
    size_t hbm_budget;      /* device bytes the admitted runs may hold together */
    int max_active;         /* admitted runs at most */
    int steps_per_quantum;  /* steps per round of the largest running run */
    int use_graph;          /* 1: each step is one graph launch */
    int device;
} hd_server_config;

enum { HD_RUN_QUEUED = 0, HD_RUN_RUNNING = 1, HD_RUN_DONE = 2, HD_RUN_FAILED = 3, HD_RUN_CANCELLED = 4 };

int hd_server_create(const hd_server_config *cfg, hd_server **out);
/* Copies the inputs; the run goes from ops->t_start until t >= (int)ops->Tmax. */
int hd_server_submit(hd_server *s, const hd_mesh *mesh, const hd_coeffs *coeffs, const hd_options *ops,
                     const hd_config *cfg, int64_t *id);
int hd_server_poll(hd_server *s, int64_t id, int *state, double *t);
/* Waits until the run ends (timeout_s < 0: no limit); HD_E_STATE on timeout. */
int hd_server_wait(hd_server *s, int64_t id, double timeout_s, int *state);
/* Final U, V, Z, FS, DH or KHOUOT (as doubles) of a finished run, dense
 * (N+3) x (M+3). */
int hd_server_result(hd_server *s, int64_t id, int field, double *dst, size_t n);
int hd_server_error(hd_server *s, int64_t id, char *buf, size_t n);
/* A queued run ends as cancelled at once (also while every slot is
 * taken); a running one at its context's next round boundary, or, in a
 * batched context, when that context ends. */
int hd_server_cancel(hd_server *s, int64_t id);
/* Drops a finished run's stored results. */
int hd_server_release(hd_server *s, int64_t id);
/* JSON: counts (active: admitted runs not yet ended; batched: runs
 * admitted into a batched context of two or more), p50/p99/mean
 * completion latency (s, submit to results on the host), cell-steps of
 * completed runs and their rate.  A cancelled run of a batched context
 * ends when its context does. */
int hd_server_stats(hd_server *s, char *buf, size_t n);
int hd_server_destroy(hd_server *s);

#ifdef __cplusplus
}
#endif

#endif /* HYDIST_H */
