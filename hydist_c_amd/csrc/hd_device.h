// hd_device.h - device-side data layout and per-cell arithmetic of the
// Hydist_C ADI timestep on gfx950.
//
// Every expression below is evaluated in the reference's operand order and
// the library is built with -ffp-contract=off, so values formed from
// + - * / sqrt are bitwise those of the reference's FP64 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hydist.h"

#define HD_DEV __device__ __forceinline__

namespace hd {

// One scenario's device buffers.  2-D fields are row-major [i][j] with
// pitch P (elements, a multiple of 32 so every row starts 256-B aligned);
// rows 0..N+2 are meaningful, row N+3 is spare.
struct DevState {
    int M, N, P, H;        // H = allocated rows (N + 4)
    int capH, capV;        // per-line capacity of the segment tables
    int total_time;
    int btype[4], bienQ[4];
    int hasCC[4];
    int quirks, ketdinh;
    hd_coeffs c;
    // reciprocals of c's spacing constants (1/dXbp, 1/dYbp, 1/dX2, 1/dY2),
    // for the fast forms' assembly, which multiplies where the reference
    // divides; set from c with every descriptor upload
    double r_dXbp, r_dYbp, r_dX2, r_dY2;
    double Dxr_m03;  // pow(c.Dxr, -0.3) of the non-cohesive source and of Qb
    // the other constant powers of Calculate_Qb: pow((Sx - 1) g, 0.5) and
    // pow(dm, 1.5) (the same values as the reference forms per cell)
    double qb_sg, qb_dm;
    // inputs
    const double *h, *hsnham, *visc, *Fw, *Tsxw, *Tsyw;
    const double *bc[4], *CC[4];
    // derived once at init
    double *htaiz, *htaiz_bd, *Kx1, *Ky1, *hi, *hmax;
    // state
    double *u, *v, *z, *t_u, *t_v, *t_z, *H_moi, *Htdu, *Htdv, *FS, *tFS, *VTH, *Kx, *Ky, *Qbx,
        *Qby, *dH;
    int *khouot;
    double *ubt, *ubp, *vbt, *vbd;
    // wet-segment tables: rows (along j) and columns (along i)
    int *segH_n, *segH_f, *segH_l;
    int *segV_n, *segV_f, *segV_l;
    // per cell (layout of the fields): index of the row segment (segidH)
    // and of the column segment (segidV) containing the cell, -1 outside
    // every segment; written with the tables by the segmentation kernel
    int *segidH, *segidV;
    // per cell, two ints: the first and last row of the column segment
    // containing the cell (-1, -1 outside), so the column kernels, whose
    // lanes run across columns, read a segment's bounds coalesced instead
    // of from 64 different lines' tables
    int *segflV;
    // scratch
    double *tmp;       // smoothing output (full field)
    // the line-solver tuner's copies of the fields a row-line sweep writes
    // (t_z and t_v, or tFS), kept while it times whole sweeps
    double *keep0, *keep1;
    double *acc;       // [0] boundary inflow volume, [1] mass-balance result
    // re-wetting candidates: bit (i & 63) of word [(i >> 6) * P + j] marks
    // row i of column j; cleared by the re-wetting pass that reads it
    unsigned long long *rwmask;
    // simulation time of the current half-step (one value shared by every
    // scenario of a context), set or advanced by k_clock
    double *clock;
};

HD_DEV int at(const DevState &s, int i, int j) { return i * s.P + j; }

// ---------------------------------------------------------------------------
// Coefficients of one cell of an implicit hydraulic line
// (_uzSolver_calculate_preindex, UVZSolver_multithread.cpp:210-278;
//  _vzSolver_calculate_preindex, :142-208).
// ---------------------------------------------------------------------------
struct Pre {
    double f1, f2, f3, f5, a2, c2, d2;
};

HD_DEV Pre uz_preindex(const DevState &s, int i, int j) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Htdu = s.Htdu, *Htdv = s.Htdv, *H_moi = s.H_moi;
    const double *VIS = s.visc, *Kx1 = s.Kx1, *Tsxw = s.Tsxw;
    const double dT = c.dT, dXbp = c.dXbp, dYbp = c.dYbp, dY2 = c.dY2, dTchia2dY = c.dTchia2dY,
                 dTchia2dX = c.dTchia2dX, H_TINH = c.H_TINH;
    Pre r;
    double p = 0.0, q = 0.0;
    const int ij = i * W + j;
    const double uij = u[ij];
    double vtb = (v[i * W + j - 1] + v[ij] + v[(i + 1) * W + j - 1] + v[(i + 1) * W + j]) * 0.25;
    r.f1 = dTchia2dX * uij + VIS[ij] * dT / dXbp;
    r.f2 = -(2 + Kx1[ij] * dT * sqrt(uij * uij + vtb * vtb) / Htdu[ij] + (2 * dT * VIS[ij]) / dXbp);
    r.f3 = dT * VIS[ij] / dXbp - dTchia2dX * uij;
    if (H_moi[ij - 1] <= H_TINH) {
        if (vtb < 0) {
            p = vtb * (-3 * uij + 4 * u[ij + 1] - u[ij + 2]) / dY2;
            q = (uij - 2 * u[ij + 1] + u[ij + 2]) / dYbp;
        }
    } else {
        if (H_moi[ij + 1] <= H_TINH) {
            if ((H_moi[ij - 2] > H_TINH) && (vtb > 0)) {
                p = vtb * (3 * uij - 4 * u[ij - 1] + u[ij - 2]) / dY2;
                q = (uij - 2 * u[ij - 1] + u[ij - 2]) / dYbp;
            }
        } else {
            p = vtb * (u[ij + 1] - u[ij - 1]) / dY2;
            q = (u[ij + 1] - 2 * uij + u[ij - 1]) / dYbp;
        }
    }
    r.f5 = -2 * uij + dT * p - dT * c.CORIOLIS_FORCE * vtb - dT * VIS[ij] * q -
           dT * (c.Windx - Tsxw[ij]) / Htdu[ij];
    r.c2 = dTchia2dX * Htdu[ij];
    r.a2 = -dTchia2dX * Htdu[(i - 1) * W + j];
    r.d2 = z[ij] - dTchia2dY * (Htdv[ij] * v[ij] - Htdv[ij - 1] * v[ij - 1]);
    return r;
}

HD_DEV Pre vz_preindex(const DevState &s, int i, int j) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Htdu = s.Htdu, *Htdv = s.Htdv, *H_moi = s.H_moi;
    const double *VIS = s.visc, *Ky1 = s.Ky1, *Tsyw = s.Tsyw;
    const double dT = c.dT, dXbp = c.dXbp, dYbp = c.dYbp, dX2 = c.dX2, dTchia2dY = c.dTchia2dY,
                 dTchia2dX = c.dTchia2dX, H_TINH = c.H_TINH;
    Pre r;
    double p = 0.0, q = 0.0;
    const int ij = i * W + j;
    const double vij = v[ij];
    double utb = (u[(i - 1) * W + j] + u[ij] + u[(i - 1) * W + j + 1] + u[ij + 1]) * 0.25;
    r.f1 = dTchia2dY * vij + VIS[ij] * dT / dYbp;
    r.f2 = -(2 + Ky1[ij] * dT * sqrt(vij * vij + utb * utb) / Htdv[ij] + (2 * dT * VIS[ij]) / dYbp);
    r.f3 = dT * VIS[ij] / dYbp - dTchia2dY * vij;
    if (H_moi[(i - 1) * W + j] <= H_TINH) {
        if (utb < 0) {
            q = utb * (-3 * vij + 4 * v[(i + 1) * W + j] - v[(i + 2) * W + j]) / dX2;
            p = (vij - 2 * v[(i + 1) * W + j] + v[(i + 2) * W + j]) / dXbp;
        }
    } else {
        if (H_moi[(i + 1) * W + j] <= H_TINH) {
            if ((H_moi[(i - 2) * W + j] > H_TINH) && (utb > 0)) {
                q = utb * (3 * vij - 4 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / dX2;
                p = (vij - 2 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / dXbp;
            }
        } else {
            q = utb * (v[(i + 1) * W + j] - v[(i - 1) * W + j]) / dX2;
            p = (v[(i + 1) * W + j] - 2 * vij + v[(i - 1) * W + j]) / dXbp;
        }
    }
    r.f5 = -2 * vij + dT * q + dT * c.CORIOLIS_FORCE * utb - dT * VIS[ij] * p -
           dT * (c.Windy - Tsyw[ij]) / Htdv[ij];
    r.c2 = dTchia2dY * Htdv[ij];
    r.a2 = -dTchia2dY * Htdv[ij - 1];
    r.d2 = z[ij] - dTchia2dX * (Htdu[ij] * u[ij] - Htdu[(i - 1) * W + j] * u[(i - 1) * W + j]);
    return r;
}

// The fast forms' preindex: the same terms with every division by a
// spacing constant replaced by a product with its reciprocal and the two
// divisions by the face depth by one reciprocal (results differ from the
// reference's by rounding only).
HD_DEV Pre uz_preindex_fast(const DevState &s, int i, int j) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Htdu = s.Htdu, *Htdv = s.Htdv, *H_moi = s.H_moi;
    const double *VIS = s.visc, *Kx1 = s.Kx1, *Tsxw = s.Tsxw;
    const double dT = c.dT, dTchia2dY = c.dTchia2dY, dTchia2dX = c.dTchia2dX, H_TINH = c.H_TINH;
    const double rXbp = s.r_dXbp, rYbp = s.r_dYbp, rY2 = s.r_dY2;
    Pre r;
    double p = 0.0, q = 0.0;
    const int ij = i * W + j;
    const double uij = u[ij], vis = VIS[ij];
    const double rH = 1.0 / Htdu[ij];
    double vtb = (v[i * W + j - 1] + v[ij] + v[(i + 1) * W + j - 1] + v[(i + 1) * W + j]) * 0.25;
    r.f1 = dTchia2dX * uij + vis * dT * rXbp;
    r.f2 = -(2 + Kx1[ij] * dT * sqrt(uij * uij + vtb * vtb) * rH + (2 * dT * vis) * rXbp);
    r.f3 = dT * vis * rXbp - dTchia2dX * uij;
    if (H_moi[ij - 1] <= H_TINH) {
        if (vtb < 0) {
            p = vtb * (-3 * uij + 4 * u[ij + 1] - u[ij + 2]) * rY2;
            q = (uij - 2 * u[ij + 1] + u[ij + 2]) * rYbp;
        }
    } else {
        if (H_moi[ij + 1] <= H_TINH) {
            if ((H_moi[ij - 2] > H_TINH) && (vtb > 0)) {
                p = vtb * (3 * uij - 4 * u[ij - 1] + u[ij - 2]) * rY2;
                q = (uij - 2 * u[ij - 1] + u[ij - 2]) * rYbp;
            }
        } else {
            p = vtb * (u[ij + 1] - u[ij - 1]) * rY2;
            q = (u[ij + 1] - 2 * uij + u[ij - 1]) * rYbp;
        }
    }
    r.f5 = -2 * uij + dT * p - dT * c.CORIOLIS_FORCE * vtb - dT * vis * q - dT * (c.Windx - Tsxw[ij]) * rH;
    r.c2 = dTchia2dX * Htdu[ij];
    r.a2 = -dTchia2dX * Htdu[(i - 1) * W + j];
    r.d2 = z[ij] - dTchia2dY * (Htdv[ij] * v[ij] - Htdv[ij - 1] * v[ij - 1]);
    return r;
}

HD_DEV Pre vz_preindex_fast(const DevState &s, int i, int j) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Htdu = s.Htdu, *Htdv = s.Htdv, *H_moi = s.H_moi;
    const double *VIS = s.visc, *Ky1 = s.Ky1, *Tsyw = s.Tsyw;
    const double dT = c.dT, dTchia2dY = c.dTchia2dY, dTchia2dX = c.dTchia2dX, H_TINH = c.H_TINH;
    const double rXbp = s.r_dXbp, rYbp = s.r_dYbp, rX2 = s.r_dX2;
    Pre r;
    double p = 0.0, q = 0.0;
    const int ij = i * W + j;
    const double vij = v[ij], vis = VIS[ij];
    const double rH = 1.0 / Htdv[ij];
    double utb = (u[(i - 1) * W + j] + u[ij] + u[(i - 1) * W + j + 1] + u[ij + 1]) * 0.25;
    r.f1 = dTchia2dY * vij + vis * dT * rYbp;
    r.f2 = -(2 + Ky1[ij] * dT * sqrt(vij * vij + utb * utb) * rH + (2 * dT * vis) * rYbp);
    r.f3 = dT * vis * rYbp - dTchia2dY * vij;
    if (H_moi[(i - 1) * W + j] <= H_TINH) {
        if (utb < 0) {
            q = utb * (-3 * vij + 4 * v[(i + 1) * W + j] - v[(i + 2) * W + j]) * rX2;
            p = (vij - 2 * v[(i + 1) * W + j] + v[(i + 2) * W + j]) * rXbp;
        }
    } else {
        if (H_moi[(i + 1) * W + j] <= H_TINH) {
            if ((H_moi[(i - 2) * W + j] > H_TINH) && (utb > 0)) {
                q = utb * (3 * vij - 4 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) * rX2;
                p = (vij - 2 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) * rXbp;
            }
        } else {
            q = utb * (v[(i + 1) * W + j] - v[(i - 1) * W + j]) * rX2;
            p = (v[(i + 1) * W + j] - 2 * vij + v[(i - 1) * W + j]) * rXbp;
        }
    }
    r.f5 = -2 * vij + dT * q + dT * c.CORIOLIS_FORCE * utb - dT * vis * p - dT * (c.Windy - Tsyw[ij]) * rH;
    r.c2 = dTchia2dY * Htdv[ij];
    r.a2 = -dTchia2dY * Htdv[ij - 1];
    r.d2 = z[ij] - dTchia2dX * (Htdu[ij] * u[ij] - Htdu[(i - 1) * W + j] * u[(i - 1) * W + j]);
    return r;
}

// Momentum row after continuity is substituted (_calculate_abcd,
// UVZSolver_multithread.cpp:280-342) at position p of segment [f, l]:
// interior formula, ran1 override at f, ran2 override at l-1, or the short
// form for segments of one or two cells.  `cur` is the cell's
// coefficients, `nxt` those of p+1 (only its continuity part a2, c2, d2 is
// read; unused by the short form and the ran2 override).
struct Mom {
    double a1, b1, c1, d1;
};

// The fast forms' version: one reciprocal each of cur.a2 and nxt.c2.
HD_DEV Mom calc_abcd_fast(int p, int f, int l, double f4, bool ran1, bool ran2, const Pre &cur,
                          const Pre &nxt) {
    Mom m;
    if (l - 1 > f) {
        const double ra = 1.0 / cur.a2, rc = 1.0 / nxt.c2;
        const double ka = cur.f1 * ra, kc = cur.f3 * rc;
        if (ran1 && p == f) {
            m.a1 = f4;
            m.b1 = cur.f2 - kc * nxt.a2;
            m.c1 = -f4 - kc;
            m.d1 = cur.f5 - kc * nxt.d2;
        } else if (ran2 && p == l - 1) {
            m.a1 = f4 - ka;
            m.b1 = cur.f2 - ka * cur.c2;
            m.c1 = -f4;
            m.d1 = cur.f5 - ka * cur.d2;
        } else {
            m.a1 = f4 - ka;
            m.c1 = -f4 - kc;
            m.b1 = cur.f2 - ka * cur.c2 - kc * nxt.a2;
            m.d1 = cur.f5 - ka * cur.d2 - kc * nxt.d2;
        }
    } else {
        m.a1 = f4;
        m.b1 = cur.f2;
        m.d1 = cur.f5;
        m.c1 = -f4;
    }
    return m;
}

HD_DEV Mom calc_abcd(int p, int f, int l, double f4, bool ran1, bool ran2, const Pre &cur,
                     const Pre &nxt) {
    Mom m;
    if (l - 1 > f) {
        if (ran1 && p == f) {
            m.a1 = f4;
            m.b1 = cur.f2 - (cur.f3 * nxt.a2 / nxt.c2);
            m.c1 = -f4 - (cur.f3 / nxt.c2);
            m.d1 = cur.f5 - (cur.f3 * nxt.d2 / nxt.c2);
        } else if (ran2 && p == l - 1) {
            m.a1 = f4 - cur.f1 / cur.a2;
            m.b1 = cur.f2 - (cur.f1 * cur.c2 / cur.a2);
            m.c1 = -f4;
            m.d1 = cur.f5 - (cur.f1 * cur.d2 / cur.a2);
        } else {
            m.a1 = f4 - cur.f1 / cur.a2;
            m.c1 = -f4 - (cur.f3 / nxt.c2);
            m.b1 = cur.f2 - (cur.f1 * cur.c2 / cur.a2) - (cur.f3 * nxt.a2 / nxt.c2);
            m.d1 = cur.f5 - (cur.f1 * cur.d2 / cur.a2) - (cur.f3 * nxt.d2 / nxt.c2);
        }
    } else {
        m.a1 = f4;
        m.b1 = cur.f2;
        m.d1 = cur.f5;
        m.c1 = -f4;
    }
    return m;
}

// ---------------------------------------------------------------------------
// Explicit momentum of the non-swept component
// (vSolver, UVZSolver_multithread.cpp:607-663; uSolver :724-795).
// ---------------------------------------------------------------------------
HD_DEV double v_solver(const DevState &s, const double *__restrict__ t_u, int row, int col) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Ky1 = s.Ky1, *Htdv = s.Htdv, *H_moi = s.H_moi;
    const double *VIS = s.visc, *Tsyw = s.Tsyw;
    const double H_TINH = c.H_TINH, dY = c.dY, dY2 = c.dY2, dX2 = c.dX2, dXbp = c.dXbp,
                 dYbp = c.dYbp, g = c.g, HaiChiadT = c.HaiChiadT;
    const int rc = row * W + col;
    double q = 0, p = 0, tmp = 0;
    const double vrc = v[rc];
    double utb = (u[(row - 1) * W + col] + u[rc] + u[(row - 1) * W + col + 1] + u[rc + 1]) * 0.25;
    double t_utb =
        (t_u[(row - 1) * W + col] + t_u[rc] + t_u[(row - 1) * W + col + 1] + t_u[rc + 1]) * 0.25;
    p = (v[rc + 1] - v[rc - 1]) / dY2;
    p = HaiChiadT + p + Ky1[rc] * sqrt(utb * utb + vrc * vrc) / Htdv[rc];
    if (H_moi[(row - 1) * W + col] <= H_TINH) {
        if (utb < 0) {
            q = t_utb * (-3 * vrc + 4 * v[(row + 1) * W + col] - v[(row + 2) * W + col]) / dX2;
            tmp = (vrc - 2 * v[(row + 1) * W + col] + v[(row + 2) * W + col]) / dXbp;
        }
    } else {
        if (H_moi[(row + 1) * W + col] <= H_TINH) {
            if ((H_moi[(row - 2) * W + col] > H_TINH) && (utb > 0)) {
                q = t_utb * (3 * vrc - 4 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / dX2;
                tmp = (vrc - 2 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / dXbp;
            }
        } else {
            q = t_utb * (v[(row + 1) * W + col] - v[(row - 1) * W + col]) / dX2;
            tmp = (v[(row + 1) * W + col] - 2 * vrc + v[(row - 1) * W + col]) / dXbp;
        }
    }
    q = HaiChiadT * vrc - q - c.CORIOLIS_FORCE * t_utb;
    q = (q - g * (z[rc + 1] - z[rc]) / dY + VIS[rc] * (tmp + (v[rc + 1] - 2 * vrc + v[rc - 1]) / dYbp)) +
        (c.Windy - Tsyw[rc]) / Htdv[rc];
    return q / p;
}

// uSolver takes dX from dY and Windx from Windy (UVZSolver_multithread.cpp:731,734).
HD_DEV double u_solver(const DevState &s, const double *__restrict__ t_v, int row, int col) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Kx1 = s.Kx1, *Htdu = s.Htdu, *H_moi = s.H_moi;
    const double *VIS = s.visc, *Tsxw = s.Tsxw;
    const double H_TINH = c.H_TINH, dX = c.dY, dY = c.dY, Windx = c.Windy, g = c.g,
                 HaiChiadT = c.HaiChiadT;
    const int i_j = row * W + col;
    const int i_1_j = i_j + W;
    double p = 0, q = 0, tmp = 0;
    const double uij = u[i_j];
    double vtb = (v[i_j - 1] + v[i_j] + v[i_1_j - 1] + v[i_1_j]) * 0.25;
    double t_vtb = (t_v[i_j - 1] + t_v[i_j] + t_v[i_1_j - 1] + t_v[i_1_j]) * 0.25;
    p = (u[i_1_j] - u[(row - 1) * W + col]) / (dX * 2);
    p = (HaiChiadT + p + Kx1[i_j] * sqrt(vtb * vtb + uij * uij) / Htdu[i_j]);
    if (H_moi[i_j - 1] <= H_TINH) {
        if (vtb < 0) {
            q = t_vtb * (-3 * uij + 4 * u[i_j + 1] - u[i_j + 2]) / (dY * 2);
            tmp = (uij - 2 * u[i_j + 1] + u[i_j + 2]) / (dY * dY);
        }
    } else {
        if (H_moi[i_j + 1] <= H_TINH) {
            if ((H_moi[i_j - 2] > H_TINH) && (vtb > 0)) {
                q = t_vtb * (3 * uij - 4 * u[i_j - 1] + u[i_j - 2]) / (dY * 2);
                tmp = (uij - 2 * u[i_j - 1] + u[i_j - 2]) / (dY * dY);
            }
        } else {
            q = t_vtb * (u[i_j + 1] - u[i_j - 1]) / (dY * 2);
            tmp = (u[i_j + 1] - 2 * uij + u[i_j - 1]) / (dY * dY);
        }
    }
    q = HaiChiadT * uij - q + c.CORIOLIS_FORCE * t_vtb;
    q = (q - g * (z[i_1_j] - z[i_j]) / dX +
         VIS[i_j] * (tmp + (u[i_1_j] - 2 * uij + u[(row - 1) * W + col]) / (dX * dX))) +
        (Windx - Tsxw[i_j]) / Htdu[i_j];
    return q / p;
}

// ---------------------------------------------------------------------------
// Sediment cell terms (sediment_transport.cpp:11-49).
// ---------------------------------------------------------------------------
// (1 - FS)^4 of the hindered settling velocity, wss * pow(1 - FS, 4)
// (sediment_transport.cpp:94, :243, :524), as two squarings: within two
// ulp of pow and a few times cheaper than the general FP64 pow.  FS and dH
// are compared with the oracle to relative tolerances in any case (the
// device's pow and log differ from the host's in the last bit).
HD_DEV double pow4(double x) {
    const double x2 = x * x;
    return x2 * x2;
}

HD_DEV double source_chs(const hd_coeffs &c, double wsm, double Zf, double fs, double fw,
                         double vth, double h_moi) {
    double toee = c.Toe;
    double todd = c.Tod;
    double S = 0;
    double tob = c.ro * fw * (vth * vth);
    if (h_moi > c.ghtoe) toee = toee * (1 + c.hstoe * ((h_moi - c.ghtoe)));
    if (tob > toee)
        S = c.Mbochat / c.ros * (tob - toee) / toee;
    else {
        if (tob < todd) {
            double Pd = 1 - (tob / todd);
            double beta = 1 + (Zf / (1.25 + 4.75 * pow(Pd, 2.5)));
            double Cb = beta * fs;
            S = -wsm * Cb * Pd;
        }
    }
    return S;
}

HD_DEV double source_nchs(const hd_coeffs &c, double wsm, double Zf, double Uf, double h_moi,
                          double fs) {
    double gamac = 0.434 + 5.975 * Zf + 2.888 * Zf * Zf;
    double Tx = fmax(0.0, (Uf * Uf - c.Ufr * c.Ufr) / (c.Ufr * c.Ufr));
    double Cac = 0.015 * c.dm * pow(Tx, 1.5) * pow(c.Dxr, (-0.3)) / (0.05 * h_moi);
    return wsm * (Cac - gamac * fs);
}

struct FSCell {
    double wsm, Zf, S;
};

// Settling velocity, Zf and source of one cell, as formed at the top of
// _FSi/_FSj_calculate__mactrix_coeff (sediment_transport.cpp:92-108, :241-265).
HD_DEV FSCell fs_cell(const DevState &s, int pos, double t, double s_start) {
    const hd_coeffs &c = s.c;
    FSCell r;
    const double H = s.H_moi[pos], fsv = s.FS[pos], vth = s.VTH[pos];
    double cc = 18 * log(12 * H / c.Ks);
    double Uf = sqrt(c.g) * fabs(vth) / cc;
    r.wsm = c.wss * pow4(1 - fsv);
    r.Zf = 0;
    if (fsv > 0.00001) r.Zf = r.wsm / (0.4 * (Uf + 2 * r.wsm));
    if (s.ketdinh)
        r.S = source_chs(c, r.wsm, r.Zf, fsv, s.Fw[pos], vth, H);
    else
        r.S = source_nchs(c, r.wsm, r.Zf, Uf, H, fsv);
    if (t < s_start) r.S = 0;
    return r;
}

// The fast forms' cell terms: pow(Pd, 2.5) and pow(Tx, 1.5) as products
// with a square root, pow(Dxr, -0.3) precomputed, and the remaining
// divisions by one reciprocal each (rounding-level differences only).
HD_DEV double source_chs_fast(const hd_coeffs &c, double wsm, double Zf, double fs, double fw, double vth,
                              double h_moi) {
    double toee = c.Toe;
    const double todd = c.Tod;
    double S = 0;
    const double tob = c.ro * fw * (vth * vth);
    if (h_moi > c.ghtoe) toee = toee * (1 + c.hstoe * ((h_moi - c.ghtoe)));
    if (tob > toee) {
        S = c.Mbochat / c.ros * (tob - toee) / toee;
    } else if (tob < todd) {
        const double Pd = 1 - tob / todd;
        const double beta = 1 + Zf / (1.25 + 4.75 * (Pd * Pd * sqrt(Pd)));
        S = -wsm * (beta * fs) * Pd;
    }
    return S;
}

HD_DEV FSCell fs_cell_fast(const DevState &s, int pos, double t, double s_start) {
    const hd_coeffs &c = s.c;
    FSCell r;
    const double H = s.H_moi[pos], fsv = s.FS[pos], vth = s.VTH[pos];
    const double cc = 18 * log(12 * H / c.Ks);
    const double Uf = sqrt(c.g) * fabs(vth) / cc;
    r.wsm = c.wss * pow4(1 - fsv);
    r.Zf = 0;
    if (fsv > 0.00001) r.Zf = r.wsm / (0.4 * (Uf + 2 * r.wsm));
    if (s.ketdinh) {
        r.S = source_chs_fast(c, r.wsm, r.Zf, fsv, s.Fw[pos], vth, H);
    } else {
        const double gamac = 0.434 + 5.975 * r.Zf + 2.888 * r.Zf * r.Zf;
        const double Tx = fmax(0.0, (Uf * Uf - c.Ufr * c.Ufr) / (c.Ufr * c.Ufr));
        const double Cac = 0.015 * c.dm * (Tx * sqrt(Tx)) * s.Dxr_m03 / (0.05 * H);
        r.S = r.wsm * (Cac - gamac * fsv);
    }
    if (t < s_start) r.S = 0;
    return r;
}

// Wall tests of locate_segment_u / locate_segment_v
// (support_funcs.cpp:823-826, :845-851).  dir 1: line = column j; 0: row i.
HD_DEV void wall_flags(const DevState &s, int dir, int line, int first, int last, bool &ran1,
                       bool &ran2) {
    const int W = s.P;
    const double ND = s.c.NANGDAY;
    const double *h = s.h;
    if (dir == 1) {
        const int j = line;
        ran1 = first > 2;
        if (first == 2) ran1 = (h[1 * W + j] + h[1 * W + j - 1]) * 0.5 == ND;
        ran2 = last < s.N;
        if (last == s.N) ran2 = (h[s.N * W + j] + h[s.N * W + j - 1]) * 0.5 == ND;
    } else {
        const int i = line;
        ran1 = (first > 2) ||
               ((first == 2) && ((h[i * W + first - 1] + h[(i - 1) * W + first - 1]) * 0.5 == ND));
        ran2 = (last < s.M) || ((last == s.M) && ((h[i * W + last] + h[(i - 1) * W + last]) * 0.5 == ND));
    }
}

// ---------------------------------------------------------------------------
// Segment system structure (_calculate_matrix_coeff,
// UVZSolver_multithread.cpp:345-470): layout of the first row, last row
// index sn and the assembly's end type.
// ---------------------------------------------------------------------------
enum { END_WALL = 0, END_Q = 1, END_LONG = 2 };

struct SegInfo {
    int f, l, sn;
    bool ran1, ran2, isB;
    int mat_end;  // END_* used by the assembly
};

template <int DIR>
HD_DEV SegInfo seg_info(const DevState &s, int line, int f, int l, bool dkQ1, bool dkQ2, bool q11) {
    SegInfo g;
    g.f = f;
    g.l = l;
    const int L = DIR == 1 ? s.N : s.M;
    wall_flags(s, DIR, line, f, l, g.ran1, g.ran2);
    int sn = 2 * (l - f);
    if (g.ran1) {
        g.isB = true;
        if (!g.ran2) {
            if (dkQ2 && l == L) {
                g.mat_end = END_Q;
            } else {
                sn--;
                g.mat_end = END_LONG;
            }
        } else {
            g.mat_end = END_WALL;
        }
    } else {
        g.isB = dkQ1 && f == 2;
        const bool endQ = q11 ? dkQ1 : dkQ2;
        if (!g.ran2) {
            if (endQ && l == L) {
                if (!g.isB) sn--;
                g.mat_end = END_Q;
            } else {
                sn--;
                if (!g.isB) sn -= 1;
                g.mat_end = END_LONG;
            }
        } else {
            if (!g.isB) sn--;
            g.mat_end = END_WALL;
        }
    }
    g.sn = sn;
    return g;
}

template <int DIR, bool FAST = false>
HD_DEV Pre preindex(const DevState &s, int line, int p) {
    if constexpr (FAST) return DIR == 1 ? uz_preindex_fast(s, p, line) : vz_preindex_fast(s, line, p);
    return DIR == 1 ? uz_preindex(s, p, line) : vz_preindex(s, line, p);
}

// The continuity coefficients a2, c2, d2 of a cell alone (the same
// expressions as in the preindex functions above; f1..f5 are left 0): what
// the momentum row of the cell before reads from this one, and what the
// extraction's end formulas read.
template <int DIR>
HD_DEV Pre continuity(const DevState &s, int line, int p) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double *u = s.u, *v = s.v, *z = s.z, *Htdu = s.Htdu, *Htdv = s.Htdv;
    Pre r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (DIR == 1) {
        const int i = p, j = line, ij = i * W + j;
        r.c2 = c.dTchia2dX * Htdu[ij];
        r.a2 = -c.dTchia2dX * Htdu[(i - 1) * W + j];
        r.d2 = z[ij] - c.dTchia2dY * (Htdv[ij] * v[ij] - Htdv[ij - 1] * v[ij - 1]);
    } else {
        const int i = line, j = p, ij = i * W + j;
        r.c2 = c.dTchia2dY * Htdv[ij];
        r.a2 = -c.dTchia2dY * Htdv[ij - 1];
        r.d2 = z[ij] - c.dTchia2dX * (Htdu[ij] * u[ij] - Htdu[(i - 1) * W + j] * u[(i - 1) * W + j]);
    }
    return r;
}


// FS row of cell (i, j) of a column segment [first, last]
// (_FSj_calculate__mactrix_coeff, sediment_transport.cpp:204-313).
HD_DEV void fs_row_j(const DevState &s, int i, int j, int first, int last, bool ran1, bool ran2,
                     double t, double s_start, double &A, double &B, double &Cc, double &D) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double dX = c.dX, dY = c.dY, dT = c.dT, H_TINH = c.H_TINH;
    const double *H = s.H_moi, *FS = s.FS, *t_u = s.t_u, *t_v = s.t_v;
    const double *Htdu = s.Htdu, *Htdv = s.Htdv, *Kx = s.Kx, *Ky = s.Ky;
    const int pos = i * W + j;
    const double gamav = 0.98 - 0.198 * 0.0 + 0.032 * 0.0 * 0.0;  // formed before Zf (:247)
    FSCell fc = fs_cell(s, pos, t, s_start);
    A = -gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / (dX * 2) - Htdu[pos - W] * Kx[pos - W] / (H[pos] * (dX * dX));
    Cc = gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / (dX * 2) - Htdu[pos] * Kx[pos] / (H[pos] * (dX * dX));
    B = 2.0 / dT + (Htdu[pos] * Kx[pos] + Htdu[pos - W] * Kx[pos - W]) / (H[pos] * (dX * dX));
    double p = 0, q = 0;
    if ((H[pos - 1] > H_TINH) && (H[pos + 2] > H_TINH)) p = (FS[pos + 1] - FS[pos]);
    if ((H[pos - 2] > H_TINH) && (H[pos + 1] > H_TINH)) q = (FS[pos] - FS[pos - 1]);
    p = (1 / (H[pos] * (dY * dY))) * (Htdv[pos] * Ky[pos] * p - Htdv[pos - 1] * Ky[pos - 1] * q);
    if ((H[pos - 2] > H_TINH) && (H[pos + 2] > H_TINH)) {
        q = (FS[pos + 1] - FS[pos - 1]) / (dY * 2);
        q = (t_v[pos] + t_v[pos - 1]) * 0.5 * q * gamav;
    } else
        q = 0;
    D = FS[pos] / (dT * 0.5) - q + p + (fc.S / H[pos]);
    if (i == first + 1) {
        if ((ran1) || (t_u[first * W + j] == 0))
            B += A;
        else
            D -= A * FS[first * W + j];
    }
    if (i == last - 1) {
        if ((ran2) && (t_u[last * W + j] == 0))
            B += Cc;
        else
            D -= Cc * FS[last * W + j];
    }
}

// FS row of cell (i, j) of a row segment
// (_FSi_calculate__mactrix_coeff, sediment_transport.cpp:51-164).
HD_DEV void fs_row_i(const DevState &s, int i, int j, int first, int last, bool ran1, bool ran2,
                     double t, double s_start, double &A, double &B, double &Cc, double &D) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double dX = c.dX, dY = c.dY, dT = c.dT, H_TINH = c.H_TINH;
    const double *H = s.H_moi, *FS = s.FS, *t_u = s.t_u, *t_v = s.t_v;
    const double *Htdu = s.Htdu, *Htdv = s.Htdv, *Kx = s.Kx, *Ky = s.Ky;
    const int pos = i * W + j;
    FSCell fc = fs_cell(s, pos, t, s_start);
    const double gamav = 0.98 - 0.198 * fc.Zf + 0.032 * fc.Zf * fc.Zf;
    A = -gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / (dY * 2) - Htdv[pos - 1] * Ky[pos - 1] / (H[pos] * (dY * dY));
    Cc = gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / (dY * 2) - Htdv[pos] * Ky[pos] / (H[pos] * (dY * dY));
    B = 2.0 / dT + (Htdv[pos] * Ky[pos] + Htdv[pos - 1] * Ky[pos - 1]) / (H[pos] * (dY * dY));
    double p = 0, q = 0;
    if ((H[pos + 2 * W] > H_TINH) && (H[pos - W] > H_TINH)) p = (FS[pos + W] - FS[pos]);
    if (H[pos - 2 * W] > H_TINH && H[pos + W] > H_TINH) q = (FS[pos] - FS[pos - W]);
    p = (1 / (H[pos] * (dX * dX))) * (Htdu[pos] * Kx[pos] * p - Htdu[pos - W] * Kx[pos - W] * q);
    if (H[pos + 2 * W] > H_TINH && H[pos - 2 * W] > H_TINH) {
        q = (FS[pos + W] - FS[pos - W]) / (dX * 2);
        q = (t_u[pos] + t_u[pos - W]) * 0.5 * q * gamav;
    } else
        q = 0;
    D = FS[pos] / (dT * 0.5) - q + p + (fc.S / H[pos]);
    if (j == first + 1) {
        if ((ran1) || (t_v[i * W + first] == 0))
            B = B + A;
        else
            D = D - A * FS[i * W + first];
    }
    if (j == last - 1) {
        if ((ran2) || (t_v[i * W + last] == 0))
            B = B + Cc;
        else
            D = D - Cc * FS[i * W + last];
    }
}

// The fast forms' FS rows: fs_row_j / fs_row_i with fs_cell_fast and the
// divisions by the depth and the spacing constants replaced by products
// with reciprocals (rounding-level differences only).
HD_DEV void fs_row_j_fast(const DevState &s, int i, int j, int first, int last, bool ran1, bool ran2,
                          double t, double s_start, double &A, double &B, double &Cc, double &D) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double H_TINH = c.H_TINH;
    const double *H = s.H_moi, *FS = s.FS, *t_u = s.t_u, *t_v = s.t_v;
    const double *Htdu = s.Htdu, *Htdv = s.Htdv, *Kx = s.Kx, *Ky = s.Ky;
    const int pos = i * W + j;
    const double gamav = 0.98;  // formed before Zf (sediment_transport.cpp:247)
    const FSCell fc = fs_cell_fast(s, pos, t, s_start);
    const double rH = 1.0 / H[pos], rdiff = rH * s.r_dXbp;
    const double adv = gamav * 0.5 * (t_u[pos] + t_u[pos - W]) * s.r_dX2;
    const double km = Htdu[pos - W] * Kx[pos - W], kp = Htdu[pos] * Kx[pos];
    A = -adv - km * rdiff;
    Cc = adv - kp * rdiff;
    B = c.HaiChiadT + (kp + km) * rdiff;
    double p = 0, q = 0;
    if ((H[pos - 1] > H_TINH) && (H[pos + 2] > H_TINH)) p = (FS[pos + 1] - FS[pos]);
    if ((H[pos - 2] > H_TINH) && (H[pos + 1] > H_TINH)) q = (FS[pos] - FS[pos - 1]);
    p = (rH * s.r_dYbp) * (Htdv[pos] * Ky[pos] * p - Htdv[pos - 1] * Ky[pos - 1] * q);
    if ((H[pos - 2] > H_TINH) && (H[pos + 2] > H_TINH))
        q = (t_v[pos] + t_v[pos - 1]) * 0.5 * ((FS[pos + 1] - FS[pos - 1]) * s.r_dY2) * gamav;
    else
        q = 0;
    D = FS[pos] * c.HaiChiadT - q + p + fc.S * rH;
    if (i == first + 1) {
        if ((ran1) || (t_u[first * W + j] == 0))
            B += A;
        else
            D -= A * FS[first * W + j];
    }
    if (i == last - 1) {
        if ((ran2) && (t_u[last * W + j] == 0))
            B += Cc;
        else
            D -= Cc * FS[last * W + j];
    }
}

HD_DEV void fs_row_i_fast(const DevState &s, int i, int j, int first, int last, bool ran1, bool ran2,
                          double t, double s_start, double &A, double &B, double &Cc, double &D) {
    const int W = s.P;
    const hd_coeffs &c = s.c;
    const double H_TINH = c.H_TINH;
    const double *H = s.H_moi, *FS = s.FS, *t_u = s.t_u, *t_v = s.t_v;
    const double *Htdu = s.Htdu, *Htdv = s.Htdv, *Kx = s.Kx, *Ky = s.Ky;
    const int pos = i * W + j;
    const FSCell fc = fs_cell_fast(s, pos, t, s_start);
    const double gamav = 0.98 - 0.198 * fc.Zf + 0.032 * fc.Zf * fc.Zf;
    const double rH = 1.0 / H[pos], rdiff = rH * s.r_dYbp;
    const double adv = gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) * s.r_dY2;
    const double km = Htdv[pos - 1] * Ky[pos - 1], kp = Htdv[pos] * Ky[pos];
    A = -adv - km * rdiff;
    Cc = adv - kp * rdiff;
    B = c.HaiChiadT + (kp + km) * rdiff;
    double p = 0, q = 0;
    if ((H[pos + 2 * W] > H_TINH) && (H[pos - W] > H_TINH)) p = (FS[pos + W] - FS[pos]);
    if (H[pos - 2 * W] > H_TINH && H[pos + W] > H_TINH) q = (FS[pos] - FS[pos - W]);
    p = (rH * s.r_dXbp) * (Htdu[pos] * Kx[pos] * p - Htdu[pos - W] * Kx[pos - W] * q);
    if (H[pos + 2 * W] > H_TINH && H[pos - 2 * W] > H_TINH)
        q = (t_u[pos] + t_u[pos - W]) * 0.5 * ((FS[pos + W] - FS[pos - W]) * s.r_dX2) * gamav;
    else
        q = 0;
    D = FS[pos] * c.HaiChiadT - q + p + fc.S * rH;
    if (j == first + 1) {
        if ((ran1) || (t_v[i * W + first] == 0))
            B = B + A;
        else
            D = D - A * FS[i * W + first];
    }
    if (j == last - 1) {
        if ((ran2) || (t_v[i * W + last] == 0))
            B = B + Cc;
        else
            D = D - Cc * FS[i * W + last];
    }
}

template <bool FAST>
HD_DEV void fs_row(int dir, const DevState &s, int i, int j, int first, int last, bool ran1, bool ran2, double t,
                   double s_start, double &A, double &B, double &Cc, double &D) {
    if constexpr (FAST) {
        if (dir == 1) fs_row_j_fast(s, i, j, first, last, ran1, ran2, t, s_start, A, B, Cc, D);
        else fs_row_i_fast(s, i, j, first, last, ran1, ran2, t, s_start, A, B, Cc, D);
    } else {
        if (dir == 1) fs_row_j(s, i, j, first, last, ran1, ran2, t, s_start, A, B, Cc, D);
        else fs_row_i(s, i, j, first, last, ran1, ran2, t, s_start, A, B, Cc, D);
    }
}


}  // namespace hd
