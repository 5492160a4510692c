// hd_sweep.h - buffers and entry of the implicit line sweeps (hd_sweep.hip).
#pragma once

#include <string>

#include "hd_device.h"

struct hd_ctx;

namespace hd {

enum { ROW_MAJOR = 0, LINE_MAJOR = 1 };

// Device buffers of one line sweep: the assembled tridiagonal rows, the
// solution and the Thomas scratch, each segment's first row and last row
// index (sn) in the line's row space, and (plain form, [rho][line]) each
// row's place in its system (0 outside any, 1 first, 2 interior, 3 last,
// 4 a one-row system).
// A batch of scenarios keeps one set per scenario, `stride` doubles (and
// `sstride` segment entries) apart.
struct SweepBufs {
    double *A, *B, *C, *D, *X, *Ap, *Bp;
    signed char *code;
    int *sys0, *sysn;
    int *segl;    // Thomas with one lane per system: the sweep's systems (index into sys0 / sysn)
    int *segcnt;  // and their number, per scenario
    int ld;          // leading dimension of the active layout
    int ld_rm[2];    // [dir]: lines per row in ROW_MAJOR
    int ld_lm[2];    // [dir]: rows per line in LINE_MAJOR
    size_t stride, sstride;

    __host__ __device__ SweepBufs scen(int k) const {
        SweepBufs b = *this;
        const size_t o = (size_t)k * stride, so = (size_t)k * sstride;
        b.A += o; b.B += o; b.C += o; b.D += o; b.X += o; b.Ap += o; b.Bp += o; b.code += o;
        b.sys0 += so; b.sysn += so; b.segl += so; b.segcnt += k;
        return b;
    }
};

int sweep_rows(const DevState &s, int dir, bool hyd);
int tune_line_solvers(hd_ctx *ctx, const DevState *S, const DevState &s, int n, std::string &report);
// nsys independent systems of m rows, ld elements apart, solved into x by
// the partition solver on ctx->stream (FP64, or FP32 + refinement when
// ctx->cfg.form is HD_FORM_FP32_REFINE).
int solve_batch(hd_ctx *ctx, const double *dl, const double *d, const double *du, const double *b, double *x, int m,
                int nsys, int ld);
int run_line_sweep(hd_ctx *ctx, const DevState *S, const DevState &s, int n, int dir, bool hyd, int start,
                   int end, double t, double s_start);

}  // namespace hd
