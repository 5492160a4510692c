// hd_sweep.h - buffers and entry of the implicit line sweeps (hd_sweep.hip).
#pragma once

#include "hd_device.h"

struct hd_ctx;

namespace hd {

enum { ROW_MAJOR = 0, LINE_MAJOR = 1 };

// Device buffers of one line sweep: the assembled tridiagonal rows, the
// solution and the Thomas scratch, plus each segment's first row and last
// row index (sn) in the line's row space.
struct SweepBufs {
    double *A, *B, *C, *D, *X, *Ap, *Bp;
    int *sys0, *sysn;
    int ld;          // leading dimension of the active layout
    int ld_rm[2];    // [dir]: lines per row in ROW_MAJOR
    int ld_lm[2];    // [dir]: rows per line in LINE_MAJOR
};

int sweep_rows(const DevState &s, int dir, bool hyd);
int run_line_sweep(hd_ctx *ctx, const DevState *S, const DevState &s, int n, int dir, bool hyd, int start,
                   int end, double t, double s_start);

}  // namespace hd
