/*
 * hydist_oracle.cpp - CPU FP64 oracle of the Hydist_C ADI shallow-water +
 * suspended-sediment timestep.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product).
 *
 * Each function restates one reference kernel, cited file:line into the
 * Hydist_C sources.  Arithmetic is written in the reference's expression
 * order and the file is compiled with -ffp-contract=off, so every value
 * built from + - * / sqrt is reproducible bit for bit by an implementation
 * that evaluates the same expressions.
 *
 * Where the reference's behaviour is undefined (uninitialised memory,
 * write races between threads, out-of-range launches) the semantics chosen
 * here are stated at the place they apply ("Define:").  Deterministic
 * quirks are kept ("Keep:").
 *
 * What pins each part (tests/test_oracle.py): the khouot dump that ships
 * with the reference (Onetime_init), Load_coeffs' constants, dense and
 * LAPACK-banded solves (tridiag), brute-force run-length encoding and the
 * edge rules (segmentation), closed forms (boundary interpolation,
 * source_chs/source_nchs), and whole-trajectory invariants (lake at rest
 * stays exactly 0, closed-basin volume, inflow balance, uniform FS stays
 * uniform, assembled systems are solved).
 * tests/test_oracle_kernels.py adds closed forms for single kernels:
 * smoothing (constant and linear fields, a hand case with a dry
 * neighbour), Find_VTH, hesoK (scaling, averaging, clamps), Calculate_Qb
 * (direction, threshold, exponent 2.1), BedLoad's divergence term, and a
 * steady channel in which surface slope balances the friction term
 * (Ky1 = g hbar^(-2 mu_mn) n^2) and continuity holds.
 * tests/test_oracle_pins.py pins the rest: the momentum rows' implicit
 * advection, viscosity and friction and their explicit cross terms in
 * every upwind branch (exact on quadratic fields, with the cross
 * velocity interpolated to the swept face), the continuity rows' flux
 * form, continuity at every cell of every solved segment for each end
 * type (extraction), the FS rows (a Fourier mode's closed-form decay
 * rate, flux form, advection along and across the line, the dry-cell
 * face rules), both branches of source_nchs, BedLoad's one-sided
 * differences beside dry cells and its FS diffusion, and hand cases of
 * Reset_states_vertical's drying and cascading re-wetting.  Each of 18
 * mistakes planted in a copy of this file (signs, indices, stencil
 * weights, dropped terms, exponents, thresholds) fails at least one of
 * those checks.
 *
 * Parallelism: OpenMP over independent lines (sweeps) and rows (stencils).
 * Every parallel loop writes disjoint data, so results do not depend on
 * the thread count.
 */
#include "hydist_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Seg {
    int first, last;
};

struct Oracle {
    int M = 0, N = 0, W = 0, H = 0;
    size_t n = 0;
    hd_coeffs c{};
    hd_options ops{};
    hd_config cfg{};
    // static inputs
    std::vector<double> h, hsnham, VISCOIDX, Fw, Tsxw, Tsyw;
    std::vector<double> bc[4], CC[4];
    int total_time = 0;
    int btype[4] = {0, 0, 0, 0};
    int bienQ[4] = {0, 0, 0, 0};
    // derived at init
    std::vector<double> htaiz, htaiz_bd, Kx1, Ky1, hi;
    double hmax[4] = {0, 0, 0, 0};
    // state
    std::vector<double> u, v, z, t_u, t_v, t_z, H_moi, Htdu, Htdv, FS, tFS, VTH, Kx, Ky, Qbx, Qby,
        dH;
    std::vector<int> khouot;
    std::vector<double> ubt, ubp, vbt, vbd;
    // wet segments: segH[i] along j for row i, segV[j] along i for column j
    std::vector<std::vector<Seg>> segH, segV;
    double inflow = 0.0;
    bool inited = false;
    std::string err;

    size_t at(int i, int j) const { return (size_t)i * W + j; }
};

int fail(Oracle *o, int code, const char *msg) {
    if (o) o->err = msg;
    return code;
}

std::vector<double> *field_ptr(Oracle *o, int f) {
    switch (f) {
        case HD_F_U: return &o->u;
        case HD_F_V: return &o->v;
        case HD_F_Z: return &o->z;
        case HD_F_TU: return &o->t_u;
        case HD_F_TV: return &o->t_v;
        case HD_F_TZ: return &o->t_z;
        case HD_F_HMOI: return &o->H_moi;
        case HD_F_HTDU: return &o->Htdu;
        case HD_F_HTDV: return &o->Htdv;
        case HD_F_FS: return &o->FS;
        case HD_F_TFS: return &o->tFS;
        case HD_F_VTH: return &o->VTH;
        case HD_F_KX: return &o->Kx;
        case HD_F_KY: return &o->Ky;
        case HD_F_QBX: return &o->Qbx;
        case HD_F_QBY: return &o->Qby;
        case HD_F_DH: return &o->dH;
        case HD_F_HTAIZ: return &o->htaiz;
        case HD_F_HTAIZ_BD: return &o->htaiz_bd;
        case HD_F_KX1: return &o->Kx1;
        case HD_F_KY1: return &o->Ky1;
        case HD_F_H: return &o->h;
        case HD_F_HSNHAM: return &o->hsnham;
        case HD_F_VISCOIDX: return &o->VISCOIDX;
        case HD_F_FW: return &o->Fw;
        default: return nullptr;
    }
}

/* ------------------------------------------------------------------ */
/* Init: Onetime_init (support_funcs.cpp:21-69)                         */
/* ------------------------------------------------------------------ */
void onetime_init(Oracle &o) {
    const int M = o.M, N = o.N, W = o.W;
    const double NANGDAY = o.c.NANGDAY, g = o.c.g, mu_mn = o.c.mu_mn;
    // Thread (i,j) marks row 0 / column 0 and cell (i+1, j+1) as 2 (border);
    // the write to row N+3 is out of range (Q3) and is dropped.
    std::fill(o.khouot.begin(), o.khouot.end(), 2);
    for (int i = 1; i <= N + 1; i++) {
        for (int j = 1; j <= M + 1; j++) {
            double s = (o.h[(i - 1) * W + j - 1] + o.h[(i - 1) * W + j] + o.h[i * W + j - 1] +
                        o.h[i * W + j]);
            if (s * 0.25 > NANGDAY) {
                o.khouot[i * W + j] = 0;
                o.H_moi[i * W + j] = 0;
            }
            if (i > N || j > M) continue;
            o.htaiz[i * W + j] = (o.h[(i - 1) * W + j - 1] + o.h[(i - 1) * W + j] +
                                  o.h[i * W + j - 1] + o.h[i * W + j]) *
                                 0.25;
            o.htaiz_bd[i * W + j] = o.htaiz[i * W + j];
            if (o.h[i * W + j - 1] + o.h[i * W + j] != 0)
                o.Kx1[i * W + j] =
                    g * std::pow((o.h[i * W + j - 1] + o.h[i * W + j]) * 0.5, -2 * mu_mn) *
                    std::pow((o.hsnham[i * W + j - 1] + o.hsnham[i * W + j]) * 0.5, 2);
            if (o.h[(i - 1) * W + j] + o.h[i * W + j] != 0)
                o.Ky1[i * W + j] =
                    g * std::pow((o.h[(i - 1) * W + j] + o.h[i * W + j]) * 0.5, -2 * mu_mn) *
                    std::pow((o.hsnham[(i - 1) * W + j] + o.hsnham[i * W + j]) * 0.5, 2);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Segmentation: the scan of Find_Calculation_limits_Horizontal         */
/* (support_funcs.cpp:260-288), identical in _Vertical (:313-341).     */
/* `wet(k)` returns khouot at line position k; `end` is M or N.        */
/* Define: no 20-segment cap (Q5).                                     */
/* ------------------------------------------------------------------ */
template <class F>
void scan_line(F wet, int end, std::vector<Seg> &out) {
    out.clear();
    int start = 2;
    while (start < end) {
        if (wet(start) != 0) {
            while (wet(start) && start < end) start++;
        }
        if (start + 1 == end) start = end;
        if (wet(start) == 0 && start + 1 < end) {
            Seg s;
            s.first = start;
            int e = start;
            while (wet(e) == 0 && e < end) e++;
            if (wet(e) != 0 && e <= end) {
                s.last = e - 1;
                start = e;
            } else {
                s.last = end;
                start = end;
            }
            out.push_back(s);
        }
    }
}

/* Interpolate_FS_ng (support_funcs.cpp:191-213).  Keep literally: the
 * integer division i_deltaX / i_dX and truncation of the row index. */
void interpolate_fs_ng(Oracle &o, int location, int offset, int sign) {
    const int M = o.M, W = o.W;
    const double dT = o.c.dT, dX = o.c.dX;
    for (int j = 1; j < M; j++) {
        if (o.t_u[location * W + j] > 0) {
            double deltaX = (o.u[location * W + j] + o.t_u[location * W + j]) * 0.25 * dT;
            // Define: a source row outside the field (i < 0 or i + 1 > N + 2
            // after truncation), or not finite, is skipped; tested on the
            // double so no out-of-range value is converted to int
            const double x = offset + sign * (deltaX / dX);
            if (!(x > -1.0 && x < o.N + 2.0)) continue;
            int i = (int)x;
            int i_deltaX = (int)std::nearbyint(deltaX);
            int i_dX = (int)std::nearbyint(dX);
            o.FS[location * W + j] =
                o.FS[i * W + j] +
                sign * (o.FS[(i + 1) * W + j] - o.FS[i * W + j]) * (deltaX / dX - i_deltaX / i_dX);
        }
    }
}

/* Interpolate_FS_d (support_funcs.cpp:215-237).  Keep: loop bound i < M
 * (Q13), no sign factor; Define: rows past N+2 are skipped. */
void interpolate_fs_d(Oracle &o, int location, int offset, int sign) {
    const int M = o.M, W = o.W;
    const double dT = o.c.dT, dY = o.c.dY;
    const int iend = std::min(M - 1, o.N + 2);
    for (int i = 1; i <= iend; i++) {
        if (o.t_v[i * W + location] > 0) {
            double deltaY = (o.v[i * W + location] + o.t_v[i * W + location]) * 0.25 * dT;
            const double y = offset + sign * (deltaY / dY);
            if (!(y > -1.0 && y < o.M + 2.0)) continue;  // Define: as in interpolate_fs_ng
            int j = (int)y;
            int j_deltaY = (int)std::nearbyint(deltaY);
            int j_dY = (int)std::nearbyint(dY);
            o.FS[i * W + location] =
                o.FS[i * W + j + 1] +
                (o.FS[i * W + j + 1] - o.FS[i * W + j]) * (deltaY / dY - j_deltaY / j_dY);
        }
    }
}

/* Find_Calculation_limits_Horizontal (:238-290) and _Vertical (:293-343),
 * including their FS boundary interpolation side effects.  Define: the
 * row-N interpolation runs before the row-2 one, then the column-M one
 * (kernel launch order). */
void find_limits(Oracle &o) {
    const int M = o.M, N = o.N, W = o.W;
    interpolate_fs_ng(o, N, N - 1, -1);
    interpolate_fs_ng(o, 2, 2, 1);
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= N; i++)
        scan_line([&](int k) { return o.khouot[i * W + k]; }, M, o.segH[i]);
    interpolate_fs_d(o, M, M - 1, -1);
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= M; j++)
        scan_line([&](int k) { return o.khouot[k * W + j]; }, N, o.segV[j]);
}

/* Htuongdoi (support_funcs.cpp:346-362). */
void htuongdoi(Oracle &o) {
    const int M = o.M, N = o.N, W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= N; i++)
        for (int j = 1; j <= M; j++) {
            o.Htdu[i * W + j] =
                (o.h[i * W + j - 1] + o.h[i * W + j] + o.z[(i + 1) * W + j] + o.z[i * W + j]) * 0.5;
            o.Htdv[i * W + j] =
                (o.h[(i - 1) * W + j] + o.h[i * W + j] + o.z[i * W + j + 1] + o.z[i * W + j]) * 0.5;
        }
}

/* preprocess_data (support_funcs.cpp:482-601): Q-boundary weights and
 * hmax_* (hi layout: up [0,N+3), down [N+3, 2N+6), left [2N+6, 2N+M+9),
 * right [2N+M+9, 2N+2M+12)). */
void preprocess_data(Oracle &o) {
    const int M = o.M, N = o.N, W = o.W;
    const double NANGDAY = o.c.NANGDAY, dX = o.c.dX, ros = o.c.ros;
    double *hi = o.hi.data();
    auto side = [&](double *hs, const std::vector<Seg> &segs, auto depth) {
        for (const Seg &s : segs) {
            double sum = 0;
            for (int i = s.first; i <= s.last; i++) {
                hs[i] = depth(i) / 2.0 - NANGDAY;
                sum += std::pow(hs[i], 5.0 / 3.0);
            }
            for (int i = s.first; i <= s.last; i++) hs[i] = std::pow(hs[i], 2.0 / 3.0) / (sum * dX);
        }
    };
    side(hi, o.segV[M], [&](int i) { return o.h[i * W + M] + o.h[(i - 1) * W + M]; });
    hi += N + 3;
    side(hi, o.segV[2], [&](int i) { return o.h[i * W + 1] + o.h[(i - 1) * W + 1]; });
    hi += N + 3;
    side(hi, o.segH[2], [&](int i) { return o.h[1 * W + i] + o.h[1 * W + i - 1]; });
    hi += M + 3;
    side(hi, o.segH[N], [&](int i) { return o.h[N * W + i] + o.h[N * W + i - 1]; });

    const double *hb = o.htaiz_bd.data();
    double hu = hb[2 * W + M], hd = hb[2 * W + 2];
    for (int i = 3; i < N; i++) {
        if (hb[i * W + M] > hu) hu = hb[i * W + M];
        if (hb[i * W + 2] > hd) hd = hb[i * W + 2];
    }
    double hl = hb[2 * W + 2], hr = hb[N * W + 2];
    for (int i = 3; i < M; i++) {
        if (hb[2 * W + i] > hl) hl = hb[2 * W + i];
        if (hb[N * W + i] > hr) hr = hb[N * W + i];
    }
    if (hu != 0) hu = 1.0 / (ros * hu);
    if (hd != 0) hd = 1.0 / (ros * hd);
    if (hl != 0) hl = 1.0 / (ros * hl);
    if (hr != 0) hr = 1.0 / (ros * hr);
    o.hmax[0] = hu;
    o.hmax[1] = hd;
    o.hmax[2] = hl;
    o.hmax[3] = hr;
}

/* ------------------------------------------------------------------ */
/* Boundary update: Update_Boundary_Value (support_funcs.cpp:682-728)  */
/* with Boundary_value (:603-640) and FS_boundary (:642-677).          */
/* Keep: only segment 0 of the boundary line receives a value (Q6).    */
/* Define: boundary types are read per side (Q7); sides are applied in */
/* the order up, down, left, right (later sides win at shared cells).  */
/* ------------------------------------------------------------------ */
void update_boundary(Oracle &o, double t_in) {
    const int M = o.M, N = o.N, W = o.W;
    // engine.cpp:22 narrows t to float
    const double t = (double)(float)t_in;
    const int t1 = (int)(t / 3600);
    const double t2 = (t - (3600.0f * t1)) / 3600.0f;
    const double t2fs = (t - (3600.0 * t1)) / 3600.0;
    const double *hi_up = o.hi.data();
    const double *hi_down = hi_up + (N + 3);
    const double *hi_left = hi_down + (N + 3);
    const double *hi_right = hi_left + (M + 3);
    const int imax = std::max(M, N);

    struct SideDef {
        int side;
        bool isU;
        int location, ext;
        const std::vector<Seg> *segs;
        const double *hi;
        std::vector<double> *barr;
    };
    SideDef defs[4] = {
        {HD_SIDE_UP, false, M, M + 1, &o.segV[M], hi_up, &o.vbt},
        {HD_SIDE_DOWN, false, 2, 1, &o.segV[2], hi_down, &o.vbd},
        {HD_SIDE_LEFT, true, 2, 1, &o.segH[2], hi_left, &o.ubt},
        {HD_SIDE_RIGHT, true, N, N + 1, &o.segH[N], hi_right, &o.ubp},
    };
    const double hm[4] = {o.hmax[0], o.hmax[1], o.hmax[2], o.hmax[3]};
    for (const SideDef &d : defs) {
        if (d.segs->empty()) continue;
        const Seg s0 = (*d.segs)[0];
        const double *bc = o.bc[d.side].data();
        const double *cc = o.CC[d.side].empty() ? nullptr : o.CC[d.side].data();
        // Define: past the end of a series its last sample is held.
        const int k0 = std::min(t1, o.total_time - 1), k1 = std::min(t1 + 1, o.total_time - 1);
        const double value = bc[k0] * (1.0f - t2) + bc[k1] * t2;
        const double fsval = cc ? cc[k0] * (1.0 - t2fs) + cc[k1] * t2fs : 0.0;
        for (int i = 2; i <= imax; i++) {
            if (!(s0.first <= i && i <= s0.last)) continue;
            if (o.btype[d.side] == 2) {
                (*d.barr)[i] = value * d.hi[i];
            } else if (d.isU) {
                o.t_z[d.location * W + i] = value;
                o.t_z[d.ext * W + i] = value;
            } else {
                o.t_z[i * W + d.location] = value;
                o.t_z[i * W + d.ext] = value;
            }
            if (o.btype[d.side] == 2) {
                int fi = d.isU ? d.location : i;
                int fj = d.isU ? i : d.location;
                o.FS[fi * W + fj] = fsval * o.htaiz_bd[fi * W + fj] * hm[d.side];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Thomas: tridiag (UVZSolver_multithread.cpp:19-43).                   */
/* ------------------------------------------------------------------ */
void thomas(int sn, const double *AA, const double *BB, const double *CC, const double *DD,
            double *x, double *Ap, double *Bp) {
    if (sn == 0) {
        x[0] = DD[0] / BB[0];
        return;
    }
    Ap[0] = -CC[0] / BB[0];
    Bp[0] = DD[0] / BB[0];
    for (int i = 1; i < sn; i++) {
        double ep = AA[i] * Ap[i - 1] + BB[i];
        Ap[i] = -CC[i] / ep;
        Bp[i] = (DD[i] - (AA[i] * Bp[i - 1])) / ep;
    }
    x[sn] = (DD[sn] - (AA[sn] * Bp[sn - 1])) / (BB[sn] + (AA[sn] * Ap[sn - 1]));
    for (int i = sn - 1; i >= 0; i--) x[i] = Bp[i] + (Ap[i] * x[i + 1]);
}

/* Wall-end tests of locate_segment_u / locate_segment_v
 * (support_funcs.cpp:810-858). dir 1 = UZ (line = column j), 0 = VZ. */
void wall_flags(const Oracle &o, int dir, int line, const Seg &s, bool &ran1, bool &ran2) {
    const int W = o.W;
    const double ND = o.c.NANGDAY;
    if (dir == 1) {
        const int j = line;
        ran1 = (s.first > 2) || ((s.first == 2) && ((o.h[1 * W + j] + o.h[1 * W + j - 1]) * 0.5 == ND));
        ran2 = (s.last < o.N) ||
               ((s.last == o.N) && ((o.h[o.N * W + j] + o.h[o.N * W + j - 1]) * 0.5 == ND));
    } else {
        const int i = line;
        ran1 = (s.first > 2) ||
               ((s.first == 2) && ((o.h[i * W + s.first - 1] + o.h[(i - 1) * W + s.first - 1]) * 0.5 == ND));
        ran2 = (s.last < o.M) ||
               ((s.last == o.M) && ((o.h[i * W + s.last] + o.h[(i - 1) * W + s.last]) * 0.5 == ND));
    }
}

/* Per-line scratch of one hydraulic sweep. Index k is the position along
 * the line (i for UZ, j for VZ); arrays have size len+3. */
struct LineWork {
    std::vector<double> f1, f2, f3, f5, a1, b1, c1, d1, a2, c2, d2;
    std::vector<double> AA, BB, CC, DD, x, Ap, Bp;
    void resize(int len) {
        size_t n1 = len + 4, n2 = 2 * len + 8;
        for (auto *v : {&f1, &f2, &f3, &f5, &a1, &b1, &c1, &d1, &a2, &c2, &d2}) v->assign(n1, 0.0);
        for (auto *v : {&AA, &BB, &CC, &DD, &x, &Ap, &Bp}) v->assign(n2, 0.0);
    }
};

/* _uzSolver_calculate_preindex (UVZSolver_multithread.cpp:210-278) at
 * (i, j) for a UZ line (column j). */
void uz_preindex(const Oracle &o, int i, int j, LineWork &w) {
    const int W = o.W;
    const hd_coeffs &c = o.c;
    const double *u = o.u.data(), *v = o.v.data(), *z = o.z.data();
    const double *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *H_moi = o.H_moi.data();
    const double *VIS = o.VISCOIDX.data(), *Kx1 = o.Kx1.data(), *Tsxw = o.Tsxw.data();
    const double dT = c.dT, dXbp = c.dXbp, dYbp = c.dYbp, dY2 = c.dY2, dTchia2dY = c.dTchia2dY,
                 dTchia2dX = c.dTchia2dX, H_TINH = c.H_TINH;
    double p = 0.0, q = 0.0;
    const int ij = i * W + j;
    double vtb = (v[i * W + j - 1] + v[ij] + v[(i + 1) * W + j - 1] + v[(i + 1) * W + j]) * 0.25;
    w.f1[i] = dTchia2dX * u[ij] + VIS[ij] * dT / dXbp;
    w.f2[i] = -(2 + Kx1[ij] * dT * std::sqrt(u[ij] * u[ij] + vtb * vtb) / Htdu[ij] +
                (2 * dT * VIS[ij]) / dXbp);
    w.f3[i] = dT * VIS[ij] / dXbp - dTchia2dX * u[ij];
    if (H_moi[ij - 1] <= H_TINH) {
        if (vtb < 0) {
            p = vtb * (-3 * u[ij] + 4 * u[ij + 1] - u[ij + 2]) / dY2;
            q = (u[ij] - 2 * u[ij + 1] + u[ij + 2]) / dYbp;
        }
    } else {
        if (H_moi[ij + 1] <= H_TINH) {
            if ((H_moi[ij - 2] > H_TINH) && (vtb > 0)) {
                p = vtb * (3 * u[ij] - 4 * u[ij - 1] + u[ij - 2]) / dY2;
                q = (u[ij] - 2 * u[ij - 1] + u[ij - 2]) / dYbp;
            }
        } else {
            p = vtb * (u[ij + 1] - u[ij - 1]) / dY2;
            q = (u[ij + 1] - 2 * u[ij] + u[ij - 1]) / dYbp;
        }
    }
    w.f5[i] = -2 * u[ij] + dT * p - dT * c.CORIOLIS_FORCE * vtb - dT * VIS[ij] * q -
              dT * (c.Windx - Tsxw[ij]) / Htdu[ij];
    w.c2[i] = dTchia2dX * Htdu[ij];
    w.a2[i] = -dTchia2dX * Htdu[(i - 1) * W + j];
    w.d2[i] = z[ij] - dTchia2dY * (Htdv[ij] * v[ij] - Htdv[ij - 1] * v[ij - 1]);
}

/* _vzSolver_calculate_preindex (UVZSolver_multithread.cpp:142-208) at
 * (i, j) for a VZ line (row i). */
void vz_preindex(const Oracle &o, int i, int j, LineWork &w) {
    const int W = o.W;
    const hd_coeffs &c = o.c;
    const double *u = o.u.data(), *v = o.v.data(), *z = o.z.data();
    const double *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *H_moi = o.H_moi.data();
    const double *VIS = o.VISCOIDX.data(), *Ky1 = o.Ky1.data(), *Tsyw = o.Tsyw.data();
    const double dT = c.dT, dXbp = c.dXbp, dYbp = c.dYbp, dX2 = c.dX2, dTchia2dY = c.dTchia2dY,
                 dTchia2dX = c.dTchia2dX, H_TINH = c.H_TINH;
    double p = 0.0, q = 0.0;
    const int ij = i * W + j;
    double utb = (u[(i - 1) * W + j] + u[ij] + u[(i - 1) * W + j + 1] + u[ij + 1]) * 0.25;
    w.f1[j] = dTchia2dY * v[ij] + VIS[ij] * dT / dYbp;
    w.f2[j] = -(2 + Ky1[ij] * dT * std::sqrt(v[ij] * v[ij] + utb * utb) / Htdv[ij] +
                (2 * dT * VIS[ij]) / dYbp);
    w.f3[j] = dT * VIS[ij] / dYbp - dTchia2dY * v[ij];
    if (H_moi[(i - 1) * W + j] <= H_TINH) {
        if (utb < 0) {
            q = utb * (-3 * v[ij] + 4 * v[(i + 1) * W + j] - v[(i + 2) * W + j]) / dX2;
            p = (v[ij] - 2 * v[(i + 1) * W + j] + v[(i + 2) * W + j]) / dXbp;
        }
    } else {
        if (H_moi[(i + 1) * W + j] <= H_TINH) {
            if ((H_moi[(i - 2) * W + j] > H_TINH) && (utb > 0)) {
                q = utb * (3 * v[ij] - 4 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / dX2;
                p = (v[ij] - 2 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / dXbp;
            }
        } else {
            q = utb * (v[(i + 1) * W + j] - v[(i - 1) * W + j]) / dX2;
            p = (v[(i + 1) * W + j] - 2 * v[ij] + v[(i - 1) * W + j]) / dXbp;
        }
    }
    w.f5[j] = -2 * v[ij] + dT * q + dT * c.CORIOLIS_FORCE * utb - dT * VIS[ij] * p -
              dT * (c.Windy - Tsyw[ij]) / Htdv[ij];
    w.c2[j] = dTchia2dY * Htdv[ij];
    w.a2[j] = -dTchia2dY * Htdv[ij - 1];
    w.d2[j] = z[ij] - dTchia2dX * (Htdu[ij] * u[ij] - Htdu[(i - 1) * W + j] * u[(i - 1) * W + j]);
}

/* _calculate_abcd (UVZSolver_multithread.cpp:280-342) for one segment.
 * Define (Q12): interior formula for k in [first, last-1] first, then the
 * end overrides.  The interior value at k = last reads coefficients of
 * cell last+1, outside the segment; it never reaches the solved system
 * and is not formed. */
void calc_abcd(const Seg &s, double f4, bool ran1, bool ran2, LineWork &w) {
    const int first = s.first, last = s.last;
    if (last - 1 > first) {
        for (int k = first; k < last; k++) {
            w.a1[k] = f4 - w.f1[k] / w.a2[k];
            w.c1[k] = -f4 - (w.f3[k] / w.c2[k + 1]);
            w.b1[k] = w.f2[k] - (w.f1[k] * w.c2[k] / w.a2[k]) - (w.f3[k] * w.a2[k + 1] / w.c2[k + 1]);
            w.d1[k] = w.f5[k] - (w.f1[k] * w.d2[k] / w.a2[k]) - (w.f3[k] * w.d2[k + 1] / w.c2[k + 1]);
        }
        if (ran1) {
            w.a1[first] = f4;
            w.b1[first] = w.f2[first] - (w.f3[first] * w.a2[first + 1] / w.c2[first + 1]);
            w.c1[first] = -f4 - (w.f3[first] / w.c2[first + 1]);
            w.d1[first] = w.f5[first] - (w.f3[first] * w.d2[first + 1] / w.c2[first + 1]);
        }
        if (ran2) {
            w.a1[last - 1] = f4 - w.f1[last - 1] / w.a2[last - 1];
            w.b1[last - 1] = w.f2[last - 1] - (w.f1[last - 1] * w.c2[last - 1] / w.a2[last - 1]);
            w.c1[last - 1] = -f4;
            w.d1[last - 1] = w.f5[last - 1] - (w.f1[last - 1] * w.d2[last - 1] / w.a2[last - 1]);
        }
    } else {
        w.a1[first] = f4;
        w.b1[first] = w.f2[first];
        w.d1[first] = w.f5[first];
        w.c1[first] = -f4;
    }
}

/* _calculate_matrix_coeff (UVZSolver_multithread.cpp:345-470) with
 * bienrandau / bienlongdau (:110-139): interleaved system of one segment,
 * rows indexed from 2*first (AA[r] is row r - 2*first).  Define (Q12):
 * the per-cell rows are written first, then the start override, then the
 * end override (so the end row wins where they coincide).
 * Returns sn (last row index).  sn < 0 happens for a one-cell segment with
 * an open start and a wall end: no unknowns (Define). */
int calc_matrix(const Seg &s, bool ran1, bool ran2, double ubt_or_vbd, double ubp_or_vbt,
                double TZ_r, double TZ_l, bool dkBienQ_1, bool dkBienQ_2, int dkfr, bool q11_literal,
                LineWork &w) {
    const int first = s.first, last = s.last;
    double *AA = w.AA.data(), *BB = w.BB.data(), *CC = w.CC.data(), *DD = w.DD.data();
    auto bienrandau = [&](int k) {
        const int r = 2 * (k - first);
        AA[r] = w.a2[k];
        AA[r + 1] = w.a1[k];
        BB[r] = 1;
        BB[r + 1] = w.b1[k];
        CC[r] = w.c2[k];
        CC[r + 1] = w.c1[k];
        DD[r] = w.d2[k];
        DD[r + 1] = w.d1[k];
    };
    auto bienlongdau = [&](int k) {
        const int r = 2 * (k - first);
        AA[r] = w.a1[k];
        BB[r] = w.b1[k];
        CC[r] = w.c1[k];
        DD[r] = w.d1[k];
        if (k < last) {  // row r+1 of k == last lies past every system row
            AA[r + 1] = w.a2[k + 1];
            BB[r + 1] = 1;
            CC[r + 1] = w.c2[k + 1];
            DD[r + 1] = w.d2[k + 1];
        }
    };
    int sn = 2 * (last - first);
    if (ran1) {
        for (int k = first; k <= last; k++) bienrandau(k);
        DD[0] = w.d2[first];
        if (!ran2) {
            if (dkBienQ_2 && last == dkfr) {
                AA[sn] = w.a2[last];
                BB[sn] = 1;
                DD[sn] = w.d2[last] - w.c2[last] * ubp_or_vbt;
            } else {
                sn--;
                AA[sn] = w.a1[last - 1];
                BB[sn] = w.b1[last - 1];
                DD[sn] = w.d1[last - 1] - w.c1[last - 1] * TZ_r;
            }
        } else {
            AA[sn] = w.a2[last];
            BB[sn] = 1;
            DD[sn] = w.d2[last];
        }
    } else {
        bool isBienran;
        if (dkBienQ_1 && first == 2) {
            for (int k = first; k <= last; k++) bienrandau(k);
            DD[0] = w.d2[first] - w.a2[first] * ubt_or_vbd;
            isBienran = true;
        } else {
            for (int k = first; k <= last; k++) bienlongdau(k);
            BB[0] = w.b1[first];
            CC[0] = w.c1[first];
            DD[0] = w.d1[first] - w.a1[first] * TZ_l;
            isBienran = false;
        }
        // Keep (Q11) only when asked: the reference tests dkBienQ_1 (the
        // start side's flag) here.  Define (default): the end side's flag,
        // as the wall-start branch above and the extraction do.
        const bool endQ = q11_literal ? dkBienQ_1 : dkBienQ_2;
        if (!ran2) {
            if (endQ && last == dkfr) {
                if (!isBienran) sn--;
                if (sn >= 0) {
                    AA[sn] = w.a2[last];
                    BB[sn] = 1;
                    DD[sn] = w.d2[last] - w.c2[last] * ubp_or_vbt;
                }
            } else {
                sn--;
                if (!isBienran) sn -= 1;
                if (sn >= 0) {
                    AA[sn] = w.a1[last - 1];
                    BB[sn] = w.b1[last - 1];
                    DD[sn] = w.d1[last - 1] - w.c1[last - 1] * TZ_r;
                }
            }
        } else {
            if (!isBienran) sn--;
            if (sn >= 0) {
                AA[sn] = w.a2[last];
                BB[sn] = 1;
                DD[sn] = w.d2[last];
            }
        }
    }
    return sn;
}

/* One hydraulic sweep line: preindex, abcd, matrix, Thomas and the
 * extraction of _uzSolver_extract_solution (:547-603) /
 * _vzSolver_extract_solution (:472-544).  dir 1 = UZ (line = column j,
 * velocity t_u), 0 = VZ (line = row i, velocity t_v).
 * Define: in the extraction, cells before `last` are written first; the
 * end cell's continuity formula then reads the new t_u[last-1]. */
void sweep_line(Oracle &o, int dir, int line, LineWork &w, int only_seg = -1, int *sn_out = nullptr) {
    const int W = o.W, M = o.M, N = o.N;
    const std::vector<Seg> &segs = dir == 1 ? o.segV[line] : o.segH[line];
    const int len = dir == 1 ? N : M;
    if ((int)w.f1.size() < len + 4) w.resize(len);
    const double f4 = dir == 1 ? 2 * o.c.g * o.c.dTchia2dX : 2 * o.c.g * o.c.dTchia2dY;
    const bool q11 = (o.cfg.quirks & HD_QUIRK_Q11_LITERAL) != 0;
    double *tz = o.t_z.data();
    double *tvel = dir == 1 ? o.t_u.data() : o.t_v.data();
    // position p along the line -> flat cell index
    auto cell = [&](int p) -> size_t { return dir == 1 ? (size_t)p * W + line : (size_t)line * W + p; };
    for (int sidx = 0; sidx < (int)segs.size(); sidx++) {
        if (only_seg >= 0 && sidx != only_seg) continue;
        const Seg s = segs[sidx];
        bool ran1, ran2;
        wall_flags(o, dir, line, s, ran1, ran2);
        for (int p = s.first; p <= s.last; p++) {
            if (dir == 1)
                uz_preindex(o, p, line, w);
            else
                vz_preindex(o, line, p, w);
        }
        calc_abcd(s, f4, ran1, ran2, w);
        int sn;
        if (dir == 1) {
            sn = calc_matrix(s, ran1, ran2, o.ubt[line], o.ubp[line], tz[cell(s.last)], tz[cell(s.first)],
                             o.bienQ[2], o.bienQ[3], N, q11, w);
        } else {
            sn = calc_matrix(s, ran1, ran2, o.vbd[line], o.vbt[line], tz[cell(s.last)], tz[cell(s.first)],
                             o.bienQ[1], o.bienQ[0], M, q11, w);
        }
        if (only_seg >= 0) {
            // inspection only: the caller reads AA..DD and x, no state is written
            if (sn >= 0) thomas(sn, w.AA.data(), w.BB.data(), w.CC.data(), w.DD.data(), w.x.data(),
                                w.Ap.data(), w.Bp.data());
            if (sn_out) *sn_out = sn;
            return;
        }
        if (sn >= 0)
            thomas(sn, w.AA.data(), w.BB.data(), w.CC.data(), w.DD.data(), w.x.data(), w.Ap.data(),
                   w.Bp.data());
        // x[r] here is the reference's x[2*first + r]
        const double *x = w.x.data();
        auto X = [&](int global) { return x[global - 2 * s.first]; };
        const bool qs = dir == 1 ? o.bienQ[2] : o.bienQ[1];
        const bool qe = dir == 1 ? o.bienQ[3] : o.bienQ[0];
        const double ub_start = dir == 1 ? o.ubt[line] : o.vbd[line];
        const double ub_end = dir == 1 ? o.ubp[line] : o.vbt[line];
        const int first = s.first, last = s.last;
        for (int p = first; p < last; p++) {
            if (ran1) {
                tz[cell(p)] = X(2 * p);
                tvel[cell(p)] = X(2 * p + 1);
                tvel[cell(first - 1)] = 0;
            } else if (qs && first == 2) {
                tz[cell(p)] = X(2 * p);
                tvel[cell(p)] = X(2 * p + 1);
                tvel[cell(first - 1)] = ub_start;
            } else {
                if (p == first) {
                    tvel[cell(first)] = X(2 * first);
                    tvel[cell(first - 1)] =
                        (w.d2[first] - tz[cell(first)] - w.c2[first] * tvel[cell(first)]) / w.a2[first];
                } else {
                    tz[cell(p)] = X(2 * p - 1);
                    tvel[cell(p)] = X(2 * p);
                }
            }
        }
        if (ran2) {
            tvel[cell(last)] = 0;
            if (sn >= 0) tz[cell(last)] = X(2 * first + sn);
        } else if (qe && last == (dir == 1 ? N : M)) {
            tvel[cell(last)] = ub_end;
            if (sn >= 0) tz[cell(last)] = X(2 * first + sn);
        } else {
            tvel[cell(last)] =
                (w.d2[last] - w.a2[last] * tvel[cell(last - 1)] - tz[cell(last)]) / w.c2[last];
        }
    }
}

/* UZ sweep over columns j in [start, end] (engine.cpp:91-104). */
void uz_sweep(Oracle &o, int start, int end) {
#pragma omp parallel
    {
        LineWork w;
        w.resize(o.N);
#pragma omp for schedule(dynamic, 8)
        for (int j = start; j <= end; j++) sweep_line(o, 1, j, w);
    }
}

/* VZ sweep over rows i in [start, end] (engine.cpp:196-208).  Define:
 * the extraction covers j <= M (the reference's guard j > endidx uses
 * endidx = N, UVZSolver_multithread.cpp:851, a launch-coverage gap like
 * Q4). */
void vz_sweep(Oracle &o, int start, int end) {
#pragma omp parallel
    {
        LineWork w;
        w.resize(o.M);
#pragma omp for schedule(dynamic, 8)
        for (int i = start; i <= end; i++) sweep_line(o, 0, i, w);
    }
}

/* Normalize + update_buffer (support_funcs.cpp:761-808): Jacobi smoothing
 * of t_u (isU) or t_v on wet-interior cells. */
void normalize(Oracle &o, bool isU, std::vector<double> &tmp) {
    const int M = o.M, N = o.N, W = o.W;
    std::vector<double> &val = isU ? o.t_u : o.t_v;
    const int closenb = isU ? W : 1;
    const int farnb = isU ? 1 : W;
    const double coeff = o.c.heso;
    tmp = val;
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= N; i++)
        for (int j = 2; j <= M; j++) {
            const int pos = i * W + j;
            if (val[pos] != 0) {
                int count = 2;
                double nb = val[pos - closenb] + val[pos + closenb];
                if (o.khouot[pos - farnb] == 0) {
                    nb += val[pos - farnb];
                    count++;
                }
                if (o.khouot[pos + farnb] == 0) {
                    nb += val[pos + farnb];
                    count++;
                }
                tmp[pos] = tmp[pos] * coeff + (1 - coeff) * nb / count;
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= N; i++)
        for (int j = 2; j <= M; j++) val[i * W + j] = tmp[i * W + j];
}

/* vSolver (UVZSolver_multithread.cpp:607-663) at (row, col). */
double v_solver(const Oracle &o, int row, int col) {
    const int W = o.W;
    const hd_coeffs &c = o.c;
    const double *u = o.u.data(), *v = o.v.data(), *t_u = o.t_u.data(), *z = o.z.data();
    const double *Ky1 = o.Ky1.data(), *Htdv = o.Htdv.data(), *H_moi = o.H_moi.data();
    const double *VIS = o.VISCOIDX.data(), *Tsyw = o.Tsyw.data();
    const double H_TINH = c.H_TINH, dY = c.dY, dY2 = c.dY2, dX2 = c.dX2, dXbp = c.dXbp,
                 dYbp = c.dYbp, g = c.g, HaiChiadT = c.HaiChiadT;
    const int rc = row * W + col;
    double q = 0, p = 0, tmp = 0;
    double utb = (u[(row - 1) * W + col] + u[rc] + u[(row - 1) * W + col + 1] + u[rc + 1]) * 0.25;
    double t_utb =
        (t_u[(row - 1) * W + col] + t_u[rc] + t_u[(row - 1) * W + col + 1] + t_u[rc + 1]) * 0.25;
    p = (v[rc + 1] - v[rc - 1]) / dY2;
    p = HaiChiadT + p + Ky1[rc] * std::sqrt(utb * utb + v[rc] * v[rc]) / Htdv[rc];
    if (H_moi[(row - 1) * W + col] <= H_TINH) {
        if (utb < 0) {
            q = t_utb * (-3 * v[rc] + 4 * v[(row + 1) * W + col] - v[(row + 2) * W + col]) / dX2;
            tmp = (v[rc] - 2 * v[(row + 1) * W + col] + v[(row + 2) * W + col]) / dXbp;
        }
    } else {
        if (H_moi[(row + 1) * W + col] <= H_TINH) {
            if ((H_moi[(row - 2) * W + col] > H_TINH) && (utb > 0)) {
                q = t_utb * (3 * v[rc] - 4 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / dX2;
                tmp = (v[rc] - 2 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / dXbp;
            }
        } else {
            q = t_utb * (v[(row + 1) * W + col] - v[(row - 1) * W + col]) / dX2;
            tmp = (v[(row + 1) * W + col] - 2 * v[rc] + v[(row - 1) * W + col]) / dXbp;
        }
    }
    q = HaiChiadT * v[rc] - q - c.CORIOLIS_FORCE * t_utb;
    q = (q - g * (z[rc + 1] - z[rc]) / dY + VIS[rc] * (tmp + (v[rc + 1] - 2 * v[rc] + v[rc - 1]) / dYbp)) +
        (c.Windy - Tsyw[rc]) / Htdv[rc];
    return q / p;
}

/* uSolver (UVZSolver_multithread.cpp:724-795) at (row, col).  Keep (Q15):
 * dX is taken from dY and Windx from Windy. */
double u_solver(const Oracle &o, int row, int col) {
    const int W = o.W;
    const hd_coeffs &c = o.c;
    const double *u = o.u.data(), *v = o.v.data(), *t_v = o.t_v.data(), *z = o.z.data();
    const double *Kx1 = o.Kx1.data(), *Htdu = o.Htdu.data(), *H_moi = o.H_moi.data();
    const double *VIS = o.VISCOIDX.data(), *Tsxw = o.Tsxw.data();
    const double H_TINH = c.H_TINH, dX = c.dY, dY = c.dY, Windx = c.Windy, g = c.g,
                 HaiChiadT = c.HaiChiadT;
    const int i_j = row * W + col;
    const int i_1_j = i_j + W;
    double p = 0, q = 0, tmp = 0;
    double vtb = (v[i_j - 1] + v[i_j] + v[i_1_j - 1] + v[i_1_j]) * 0.25;
    double t_vtb = (t_v[i_j - 1] + t_v[i_j] + t_v[i_1_j - 1] + t_v[i_1_j]) * 0.25;
    p = (u[i_1_j] - u[(row - 1) * W + col]) / (dX * 2);
    p = (HaiChiadT + p + Kx1[i_j] * std::sqrt(vtb * vtb + u[i_j] * u[i_j]) / Htdu[i_j]);
    if (H_moi[i_j - 1] <= H_TINH) {
        if (vtb < 0) {
            q = t_vtb * (-3 * u[i_j] + 4 * u[i_j + 1] - u[i_j + 2]) / (dY * 2);
            tmp = (u[i_j] - 2 * u[i_j + 1] + u[i_j + 2]) / (dY * dY);
        }
    } else {
        if (H_moi[i_j + 1] <= H_TINH) {
            if ((H_moi[i_j - 2] > H_TINH) && (vtb > 0)) {
                q = t_vtb * (3 * u[i_j] - 4 * u[i_j - 1] + u[i_j - 2]) / (dY * 2);
                tmp = (u[i_j] - 2 * u[i_j - 1] + u[i_j - 2]) / (dY * dY);
            }
        } else {
            q = t_vtb * (u[i_j + 1] - u[i_j - 1]) / (dY * 2);
            tmp = (u[i_j + 1] - 2 * u[i_j] + u[i_j - 1]) / (dY * dY);
        }
    }
    q = HaiChiadT * u[i_j] - q + c.CORIOLIS_FORCE * t_vtb;
    q = (q - g * (z[i_1_j] - z[i_j]) / dX +
         VIS[i_j] * (tmp + (u[i_1_j] - 2 * u[i_j] + u[(row - 1) * W + col]) / (dX * dX))) +
        (Windx - Tsxw[i_j]) / Htdu[i_j];
    return q / p;
}

/* solveV (UVZSolver_multithread.cpp:961-978): t_v on j in [first, last-1]
 * of every row segment, rows i in [start, end] (Q4: all rows). */
void solve_v(Oracle &o, int start, int end) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = start; i <= end; i++)
        for (const Seg &s : o.segH[i]) {
            if (s.first >= s.last) continue;
            for (int j = s.first; j < s.last; j++) o.t_v[i * o.W + j] = v_solver(o, i, j);
        }
}

/* solveU (UVZSolver_multithread.cpp:938-959): t_u on i in [first, last-1]
 * of every column segment, columns j in [start, end]. */
void solve_u(Oracle &o, int start, int end) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int j = start; j <= end; j++)
        for (const Seg &s : o.segV[j]) {
            if (s.first > s.last) continue;
            for (int i = s.first; i < s.last; i++) o.t_u[i * o.W + j] = u_solver(o, i, j);
        }
}

/* update_margin_elem_V (UVZSolver_multithread.cpp:1018-1060): per row,
 * segments in order.  Keep (Q16): wall test |face depth - NANGDAY| < 1e-16. */
void margin_v(Oracle &o, int start, int end) {
    const int W = o.W, M = o.M;
    const double ND = o.c.NANGDAY;
    double *t_v = o.t_v.data();
#pragma omp parallel for schedule(static)
    for (int i = start; i <= end; i++)
        for (const Seg &s : o.segH[i]) {
            const int first = s.first, last = s.last;
            bool ran1 = false, ran2 = false;
            if ((first > 2) ||
                ((first == 2) && (std::fabs((o.h[i * W + first - 1] + o.h[(i - 1) * W + first - 1]) * 0.5 - ND) < 1e-16)))
                ran1 = true;
            if ((last < M) ||
                ((last == M) && (std::fabs((o.h[i * W + last] + o.h[(i - 1) * W + last]) * 0.5 - ND) < 1e-16)))
                ran2 = true;
            if (ran1)
                t_v[i * W + first - 1] = 0;
            else
                t_v[i * W + first - 1] = 2 * t_v[i * W + first] - t_v[i * W + first + 1];
            if (ran2)
                t_v[i * W + last] = 0;
            else
                t_v[i * W + last] = 2 * t_v[i * W + last - 1] - t_v[i * W + last - 2];
        }
}

/* update_margin_elem_U (UVZSolver_multithread.cpp:980-1016).  Keep: the
 * start test is |h + h| * 0.5 - NANGDAY < 1e-16 (abs around the sum only). */
void margin_u(Oracle &o, int start, int end) {
    const int W = o.W, N = o.N;
    const double ND = o.c.NANGDAY;
    double *t_u = o.t_u.data();
#pragma omp parallel for schedule(static)
    for (int j = start; j <= end; j++)
        for (const Seg &s : o.segV[j]) {
            const int first = s.first, last = s.last;
            bool ran1 = false, ran2 = false;
            if ((first > 2) ||
                ((first == 2) && ((std::fabs(o.h[1 * W + j] + o.h[1 * W + j - 1]) * 0.5 - ND) < 1e-16)))
                ran1 = true;
            if ((last < N) ||
                ((last == N) && (std::fabs((o.h[N * W + j] + o.h[N * W + j - 1]) * 0.5 - ND) < 1e-16)))
                ran2 = true;
            if (ran1)
                t_u[(first - 1) * W + j] = 0;
            else
                t_u[(first - 1) * W + j] = 2 * t_u[first * W + j] - t_u[(first + 1) * W + j];
            if (ran2)
                t_u[last * W + j] = 0;
            else
                t_u[last * W + j] = 2 * t_u[(last - 1) * W + j] - t_u[(last - 2) * W + j];
        }
}

/* update_h_moi (support_funcs.cpp:71-86). */
void update_h_moi(Oracle &o) {
    const int W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= o.N; i++)
        for (int j = 1; j <= o.M; j++)
            if (o.khouot[i * W + j] == 0) o.H_moi[i * W + j] = o.htaiz[i * W + j] + o.t_z[i * W + j];
}

/* update_uvz (support_funcs.cpp:732-759).  Keep (Q18): the factors come
 * from coeffs->kenhhep*, stored into int. */
void update_uvz(Oracle &o) {
    const int W = o.W;
    const int kd = (int)o.c.kenhhepd, kn = (int)o.c.kenhhepng;
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= o.N; i++)
        for (int j = 0; j <= o.M; j++) {
            const int p = i * W + j;
            o.z[p] = o.t_z[p];
            o.u[p] = o.t_u[p] * (1 - kd);
            o.v[p] = o.t_v[p] * (1 - kn);
        }
}

/* Reset_states_vertical (support_funcs.cpp:142-188): serial along i per
 * column j in [1, M]. */
void reset_states_vertical(Oracle &o) {
    const int W = o.W, N = o.N;
    const double H_TINH = o.c.H_TINH;
    double *t_z = o.t_z.data(), *z = o.z.data(), *H_moi = o.H_moi.data(), *htaiz = o.htaiz.data();
    double *FS = o.FS.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data();
    int *kh = o.khouot.data();
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= o.M; j++) {
        for (int i = 2; i <= N; i++) {
            if (t_z[i * W + j] > z[i * W + j]) {
                if (kh[(i - 1) * W + j] == 1) {
                    t_z[(i - 1) * W + j] = t_z[i * W + j];
                    H_moi[(i - 1) * W + j] = htaiz[(i - 1) * W + j] + t_z[i * W + j];
                    kh[(i - 1) * W + j] = 0;
                    FS[(i - 1) * W + j] = FS[i * W + j];
                }
                if (kh[(i + 1) * W + j] == 1) {
                    t_z[(i + 1) * W + j] = t_z[i * W + j];
                    H_moi[(i + 1) * W + j] = htaiz[(i + 1) * W + j] + t_z[i * W + j];
                    kh[(i + 1) * W + j] = 0;
                    FS[(i + 1) * W + j] = FS[i * W + j];
                }
            }
        }
    }
    // The drying pass writes t_v[i][j-1], which belongs to column j-1; it
    // only ever writes zero, so columns stay independent.
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= o.M; j++) {
        for (int i = 2; i <= N; i++) {
            if ((kh[i * W + j] == 0) && (H_moi[i * W + j] <= H_TINH)) {
                t_u[(i - 1) * W + j] = 0;
                t_u[i * W + j] = 0;
                t_v[i * W + j - 1] = 0;
                t_v[i * W + j] = 0;
                kh[i * W + j] = 1;
                FS[i * W + j] = 0;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Sediment (sediment_transport.cpp)                                    */
/* ------------------------------------------------------------------ */

/* Find_VTH (:575-602).  Define (Q4): all i in [2,N], j in [2,M] in both
 * half-steps. */
void find_vth(Oracle &o) {
    const int W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= o.N; i++)
        for (int j = 2; j <= o.M; j++) {
            const int pos = i * W + j;
            if ((o.htaiz[pos] > o.c.NANGDAY) && (o.H_moi[pos] > o.c.H_TINH)) {
                double ut = (o.t_u[pos] + o.t_u[pos - W]) * 0.5;
                double vt = (o.t_v[pos] + o.t_v[pos - 1]) * 0.5;
                o.VTH[pos] = std::sqrt(ut * ut + vt * vt);
            } else
                o.VTH[pos] = 0;
        }
}

/* hesoK (:539-573).  Cells with Htd <= 0 keep their previous K.  Rows /
 * columns 0 have Htd == 0 (never written), so i, j >= 1 covers it. */
void heso_k(Oracle &o) {
    const int W = o.W;
    const double Ks = o.c.Ks, g = o.c.g;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= o.N; i++)
        for (int j = 1; j <= o.M; j++) {
            const int pos = i * W + j;
            if (o.Htdu[pos] > 0) {
                double Cz = 18 * std::log(12 * o.Htdu[pos] / Ks);
                double k = 5.93 * std::sqrt(g) * (o.h[pos - 1] + o.h[pos]) * 0.25 * std::fabs(o.t_u[pos]) / Cz;
                o.Kx[pos] = std::fmin(100.0, std::fmax(5.0, k));
            }
            if (o.Htdv[pos] > 0) {
                double Cz = 18 * std::log(12 * o.Htdv[pos] / Ks);
                double k = 5.93 * std::sqrt(g) * (o.h[pos - W] + o.h[pos]) * 0.25 * std::fabs(o.t_v[pos]) / Cz;
                o.Ky[pos] = std::fmin(100.0, std::fmax(5.0, k));
            }
        }
}

/* source_chs (:11-38): cohesive erosion / deposition source. */
double source_chs(const hd_coeffs &c, double wsm, double Zf, double fs, double fw, double vth,
                  double h_moi) {
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
            double beta = 1 + (Zf / (1.25 + 4.75 * std::pow(Pd, 2.5)));
            double Cb = beta * fs;
            S = -wsm * Cb * Pd;
        }
    }
    return S;
}

/* source_nchs (:41-49): non-cohesive source. */
double source_nchs(const hd_coeffs &c, double wsm, double Zf, double Dxr, double Ufr, double Uf,
                   double h_moi, double fs) {
    double gamac = 0.434 + 5.975 * Zf + 2.888 * Zf * Zf;
    double Tx = std::fmax(0.0, (Uf * Uf - Ufr * Ufr) / (Ufr * Ufr));
    double Cac = 0.015 * c.dm * std::pow(Tx, 1.5) * std::pow(Dxr, (-0.3)) / (0.05 * h_moi);
    return wsm * (Cac - gamac * fs);
}

/* Settling / source terms shared by both FS sweeps. */
struct FSCell {
    double wsm, Zf, S;
};
FSCell fs_cell(const Oracle &o, int pos, double t, double s_start) {
    const hd_coeffs &c = o.c;
    FSCell r;
    double cc = 18 * std::log(12 * o.H_moi[pos] / c.Ks);
    double Uf = std::sqrt(c.g) * std::fabs(o.VTH[pos]) / cc;
    r.wsm = c.wss * std::pow((1 - o.FS[pos]), 4);
    r.Zf = 0;
    if (o.FS[pos] > 0.00001) r.Zf = r.wsm / (0.4 * (Uf + 2 * r.wsm));
    if (o.ops.ketdinh)
        r.S = source_chs(c, r.wsm, r.Zf, o.FS[pos], o.Fw[pos], o.VTH[pos], o.H_moi[pos]);
    else
        r.S = source_nchs(c, r.wsm, r.Zf, c.Dxr, c.Ufr, Uf, o.H_moi[pos], o.FS[pos]);
    if (t < s_start) r.S = 0;
    return r;
}

/* FS sweep along i for column j: Scan_FSj (:628-640) ->
 * _FSj_calculate__mactrix_coeff (:204-313), tridiag (jump 1),
 * FSj_extract_solution (:642-651) -> _FSj_extract_solution (:316-349).
 * Keep (Q13): gamav is formed before Zf (so 0.98); the end test uses &&.
 * Define: segments shorter than 3 cells have no FS system. */
void fs_line_j(Oracle &o, int j, double t, double s_start, LineWork &w, int only_seg = -1, int *sn_out = nullptr) {
    const int W = o.W;
    const hd_coeffs &c = o.c;
    const double dX = c.dX, dY = c.dY, dT = c.dT, H_TINH = c.H_TINH;
    const double *H = o.H_moi.data(), *FS = o.FS.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data();
    const double *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *Kx = o.Kx.data(), *Ky = o.Ky.data();
    for (int sidx = 0; sidx < (int)o.segV[j].size(); sidx++) {
        if (only_seg >= 0 && sidx != only_seg) continue;
        const Seg &s = o.segV[j][sidx];
        const int first = s.first, last = s.last;
        if (sn_out) *sn_out = -1;
        if (last < first + 1 || last - first < 2) continue;
        bool ran1, ran2;
        wall_flags(o, 1, j, s, ran1, ran2);
        double *AA = w.AA.data(), *BB = w.BB.data(), *CC = w.CC.data(), *DD = w.DD.data();
        for (int i = first + 1; i <= last - 1; i++) {
            const int pos = i * W + j;
            const int k = i - (first + 1);
            const double gamav = 0.98 - 0.198 * 0.0 + 0.032 * 0.0 * 0.0;
            FSCell fc = fs_cell(o, pos, t, s_start);
            AA[k] = -gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / (dX * 2) -
                    Htdu[pos - W] * Kx[pos - W] / (H[pos] * (dX * dX));
            CC[k] = gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / (dX * 2) -
                    Htdu[pos] * Kx[pos] / (H[pos] * (dX * dX));
            BB[k] = 2.0 / dT + (Htdu[pos] * Kx[pos] + Htdu[pos - W] * Kx[pos - W]) / (H[pos] * (dX * dX));
            double p = 0, q = 0;
            if ((H[pos - 1] > H_TINH) && (H[pos + 2] > H_TINH)) p = (FS[pos + 1] - FS[pos]);
            if ((H[pos - 2] > H_TINH) && (H[pos + 1] > H_TINH)) q = (FS[pos] - FS[pos - 1]);
            p = (1 / (H[pos] * (dY * dY))) * (Htdv[pos] * Ky[pos] * p - Htdv[pos - 1] * Ky[pos - 1] * q);
            if ((H[pos - 2] > H_TINH) && (H[pos + 2] > H_TINH)) {
                q = (FS[pos + 1] - FS[pos - 1]) / (dY * 2);
                q = (t_v[pos] + t_v[pos - 1]) * 0.5 * q * gamav;
            } else
                q = 0;
            DD[k] = FS[pos] / (dT * 0.5) - q + p + (fc.S / H[pos]);
            if (i == first + 1) {
                if ((ran1) || (t_u[first * W + j] == 0))
                    BB[k] += AA[k];
                else
                    DD[k] -= AA[k] * FS[first * W + j];
            }
            if (i == last - 1) {
                if ((ran2) && (t_u[last * W + j] == 0))
                    BB[k] += CC[k];
                else
                    DD[k] -= CC[k] * FS[last * W + j];
            }
        }
        const int sn = last - first - 2;
        if (only_seg >= 0) {  // inspection only: the rows, nothing solved or written
            if (sn_out) *sn_out = sn;
            return;
        }
        thomas(sn, AA, BB, CC, DD, w.x.data(), w.Ap.data(), w.Bp.data());
        const double *x = w.x.data();
        double *tFS = o.tFS.data();
        for (int i = first + 1; i <= last - 1; i++) {
            const int pos = i * W + j;
            const double xi = x[i - (first + 1)];
            tFS[pos] = xi;
            if (xi < 0) tFS[pos] = c.NDnen;
            if (i == first + 1) {
                if ((ran1) || (t_u[pos - W] == 0))
                    tFS[first * W + j] = tFS[pos];
                else
                    tFS[first * W + j] = FS[first * W + j];
            }
            if (i == last - 1) {
                if ((ran2) || (t_u[pos + W] == 0))
                    tFS[last * W + j] = tFS[pos];
                else
                    tFS[last * W + j] = FS[last * W + j];
            }
        }
    }
}

/* FS sweep along j for row i: Scan_FSi (:604-614) ->
 * _FSi_calculate__mactrix_coeff (:51-164), tridiag, FSi_extract_solution
 * (:615-625) -> _FSi_extract_solution (:167-201). */
void fs_line_i(Oracle &o, int i, double t, double s_start, LineWork &w, int only_seg = -1, int *sn_out = nullptr) {
    const int W = o.W;
    const hd_coeffs &c = o.c;
    const double dX = c.dX, dY = c.dY, dT = c.dT, H_TINH = c.H_TINH;
    const double *H = o.H_moi.data(), *FS = o.FS.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data();
    const double *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *Kx = o.Kx.data(), *Ky = o.Ky.data();
    for (int sidx = 0; sidx < (int)o.segH[i].size(); sidx++) {
        if (only_seg >= 0 && sidx != only_seg) continue;
        const Seg &s = o.segH[i][sidx];
        const int first = s.first, last = s.last;
        if (sn_out) *sn_out = -1;
        if (last < first + 1 || last - first < 2) continue;
        bool ran1, ran2;
        wall_flags(o, 0, i, s, ran1, ran2);
        double *AA = w.AA.data(), *BB = w.BB.data(), *CC = w.CC.data(), *DD = w.DD.data();
        for (int j = first + 1; j <= last - 1; j++) {
            const int pos = i * W + j;
            const int k = j - (first + 1);
            FSCell fc = fs_cell(o, pos, t, s_start);
            const double gamav = 0.98 - 0.198 * fc.Zf + 0.032 * fc.Zf * fc.Zf;
            AA[k] = -gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / (dY * 2) -
                    Htdv[pos - 1] * Ky[pos - 1] / (H[pos] * (dY * dY));
            CC[k] = gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / (dY * 2) -
                    Htdv[pos] * Ky[pos] / (H[pos] * (dY * dY));
            BB[k] = 2.0 / dT + (Htdv[pos] * Ky[pos] + Htdv[pos - 1] * Ky[pos - 1]) / (H[pos] * (dY * dY));
            double p = 0, q = 0;
            if ((H[pos + 2 * W] > H_TINH) && (H[pos - W] > H_TINH)) p = (FS[pos + W] - FS[pos]);
            if (H[pos - 2 * W] > H_TINH && H[pos + W] > H_TINH) q = (FS[pos] - FS[pos - W]);
            p = (1 / (H[pos] * (dX * dX))) * (Htdu[pos] * Kx[pos] * p - Htdu[pos - W] * Kx[pos - W] * q);
            if (H[pos + 2 * W] > H_TINH && H[pos - 2 * W] > H_TINH) {
                q = (FS[pos + W] - FS[pos - W]) / (dX * 2);
                q = (t_u[pos] + t_u[pos - W]) * 0.5 * q * gamav;
            } else
                q = 0;
            DD[k] = FS[pos] / (dT * 0.5) - q + p + (fc.S / H[pos]);
            if (j == first + 1) {
                if ((ran1) || (t_v[i * W + first] == 0))
                    BB[k] = BB[k] + AA[k];
                else
                    DD[k] = DD[k] - AA[k] * FS[i * W + first];
            }
            if (j == last - 1) {
                if ((ran2) || (t_v[i * W + last] == 0))
                    BB[k] = BB[k] + CC[k];
                else
                    DD[k] = DD[k] - CC[k] * FS[i * W + last];
            }
        }
        const int sn = last - first - 2;
        if (only_seg >= 0) {  // inspection only: the rows, nothing solved or written
            if (sn_out) *sn_out = sn;
            return;
        }
        thomas(sn, AA, BB, CC, DD, w.x.data(), w.Ap.data(), w.Bp.data());
        double *x = w.x.data();
        double *tFS = o.tFS.data();
        for (int j = first + 1; j <= last - 1; j++) {
            const int pos = i * W + j;
            double &xj = x[j - (first + 1)];
            if (xj < 0) xj = c.NDnen;
            tFS[pos] = xj;
            if (j == first + 1) {
                if ((ran1) || (t_v[i * W + first] == 0))
                    tFS[i * W + first] = tFS[pos];
                else
                    tFS[i * W + first] = FS[i * W + first];
            }
            if (j == last - 1) {
                if ((ran2) || (t_v[i * W + last] == 0))
                    tFS[i * W + last] = tFS[pos];
                else
                    tFS[i * W + last] = FS[i * W + last];
            }
        }
    }
}

/* Update_FS (:654-663). */
void update_fs(Oracle &o) {
    const int W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= o.N; i++)
        for (int j = 1; j <= o.M; j++) o.FS[i * W + j] = o.tFS[i * W + j];
}

void fs_sweep_j(Oracle &o, double t, int start, int end) {
#pragma omp parallel
    {
        LineWork w;
        w.resize(o.N);
#pragma omp for schedule(dynamic, 8)
        for (int j = start; j <= end; j++) fs_line_j(o, j, t, o.ops.bed_change_start, w);
    }
}

void fs_sweep_i(Oracle &o, double t, int start, int end) {
#pragma omp parallel
    {
        LineWork w;
        w.resize(o.M);
#pragma omp for schedule(dynamic, 8)
        for (int i = start; i <= end; i++) fs_line_i(o, i, t, o.ops.bed_change_start, w);
    }
}

/* Calculate_Qb (:352-434).  Define (Q8): Qb is formed where VTH > 0 on wet
 * cells, Toe is the coefficient, and the corner copies are braced as their
 * indentation shows. */
void calculate_qb(Oracle &o) {
    const int W = o.W, M = o.M, N = o.N;
    const hd_coeffs &c = o.c;
    double *Qbx = o.Qbx.data(), *Qby = o.Qby.data();
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= N; i++)
        for (int j = 2; j <= M; j++) {
            const int pos = i * W + j;
            double Toee = c.Toe, Tx = 0;
            if ((o.VTH[pos] > 0) && (o.khouot[pos] == 0)) {
                double Tob = c.ro * o.Fw[pos] * (o.VTH[pos] * o.VTH[pos]);
                if (o.H_moi[pos] > c.ghtoe) Toee = c.Toe * (1 + c.hstoe * ((o.H_moi[pos] - c.ghtoe)));
                if (Tob > Toee) Tx = Tob / Toee - 1;
                Qbx[pos] = 0.053 * std::pow((c.Sx - 1) * c.g, 0.5) * std::pow(c.dm, 1.5) * std::pow(Tx, 2.1) *
                           std::pow(c.Dxr, -0.3) * (o.t_u[pos] + o.t_u[pos - W]) * 0.5 / o.VTH[pos];
                Qby[pos] = 0.053 * std::pow((c.Sx - 1) * c.g, 0.5) * std::pow(c.dm, 1.5) * std::pow(Tx, 2.1) *
                           std::pow(c.Dxr, -0.3) * (o.t_v[pos] + o.t_v[pos - 1]) * 0.5 / o.VTH[pos];
            } else {
                Qbx[pos] = 0;
                Qby[pos] = 0;
            }
        }
    // ghost copies (read only interior cells written above)
    for (int i = 2; i <= N; i++)
        for (int j = 2; j <= M; j++) {
            const int pos = i * W + j;
            if (i == 2) {
                Qbx[pos - W] = Qbx[pos];
                Qby[pos - W] = Qby[pos];
            }
            if (i == N) {
                Qbx[pos + W] = Qbx[pos];
                Qby[pos + W] = Qby[pos];
            }
            if (j == 2) {
                Qbx[pos - 1] = Qbx[pos];
                Qby[pos - 1] = Qby[pos];
                if (i == 2) {
                    Qbx[1 * W + 1] = Qbx[pos];
                    Qby[1 * W + 1] = Qby[pos];
                }
                if (i == N) {
                    Qbx[(N + 1) * W + 1] = Qbx[pos];
                    Qby[(N + 1) * W + 1] = Qby[pos];
                }
            }
            if (j == M) {
                Qbx[pos + 1] = Qbx[pos];
                Qby[pos + 1] = Qby[pos];
                if (i == 2) {
                    Qbx[1 * W + M + 1] = Qbx[pos];
                    Qby[1 * W + M + 1] = Qby[pos];
                }
                if (i == N) {
                    Qbx[(N + 1) * W + M + 1] = Qbx[pos];
                    Qby[(N + 1) * W + M + 1] = Qby[pos];
                }
            }
        }
}

/* BedLoad (:664-673) -> _bed_load (:436-537) on rows i in [3, N-2] (the
 * launch's `i >= endidx` guard with endidx = N-1).  Define (Q9): the
 * downwind test is t_u[pos] > 0. */
void bed_load(Oracle &o) {
    const int W = o.W, M = o.M, N = o.N;
    const hd_coeffs &c = o.c;
    const double dX = c.dX, dY = c.dY, dT = c.dT;
    const double *Qbx = o.Qbx.data(), *Qby = o.Qby.data(), *FS = o.FS.data();
    const int *kh = o.khouot.data();
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 3; i <= N - 2; i++)
        for (const Seg &s : o.segH[i]) {
            const int first = s.first, last = s.last;
            if (last - first < 2) continue;
            for (int j = first; j <= last; j++) {
                const int pos = i * W + j;
                double p = 0, q = 0;
                if (kh[pos - W] == 1) {
                    if (o.t_u[pos] < 0) p = (-3 * Qbx[pos] + 4 * Qbx[pos + W] - Qbx[pos + 2 * W]) / (dX * 2);
                } else {
                    if (kh[pos + W] == 1) {
                        if (o.t_u[pos] > 0) p = (3 * Qbx[pos] - 4 * Qbx[pos - W] + Qbx[pos - 2 * W]) / (dX * 2);
                    } else
                        p = (Qbx[pos + W] - Qbx[pos - W]) / (dX * 2);
                }
                if (kh[pos - 1] == 1) {
                    if (o.t_v[pos] < 0) q = (-3 * Qby[pos] + 4 * Qby[pos + 1] - Qby[pos + 2]) / (dY * 2);
                } else {
                    if (kh[pos + 1] == 1) {
                        if (o.t_v[pos] > 0) q = (3 * Qby[pos] - 4 * Qby[pos - 1] + Qby[pos - 2]) / (dY * 2);
                    } else
                        q = (Qby[pos + 1] - Qby[pos - 1]) / (dY * 2);
                }
                double tH = p + q;
                p = FS[pos + W] - FS[pos];
                if ((kh[pos + 2 * W] == 1 || kh[pos - W] == 1) && (i < N)) p = 0;
                q = FS[pos] - FS[pos - W];
                if ((i > 2) && (kh[pos - 2 * W] == 1 || kh[pos + W] == 1)) q = 0;
                p = 1 / (dX * dX) * (o.Htdu[pos] * o.Kx[pos] * p - o.Htdu[pos - W] * o.Kx[pos - W] * q);
                tH = tH + p;
                p = FS[pos + 1] - FS[pos];
                if (((last < M) && (kh[pos - 1] == 1 || kh[pos + 2] == 1)) || (last == M && (kh[pos - 1] == 1)))
                    p = 0;
                q = FS[pos] - FS[pos - 1];
                if (((first > 2) && (kh[pos - 2] == 1 || kh[pos + 1] == 1)) || (first == 2 && kh[pos + 1] == 1))
                    q = 0;
                p = 1 / ((dY * dY)) * (o.Htdv[pos] * o.Ky[pos] * p - o.Htdv[pos - 1] * o.Ky[pos - 1] * q);
                tH = tH + p;
                double cc = 18 * std::log(12 * o.H_moi[pos] / c.Ks);
                double Uf = std::sqrt(c.g) * std::fabs(o.VTH[pos]) / cc;
                double wsm = c.wss * std::pow((1 - FS[pos]), 4);
                double Zf = 0, S = 0;
                if (FS[pos] > 0.0) Zf = wsm / (0.4 * (Uf + 2 * wsm));
                if (o.ops.ketdinh)
                    S = source_chs(c, wsm, Zf, FS[pos], o.Fw[pos], o.VTH[pos], o.H_moi[pos]);
                else
                    S = source_nchs(c, wsm, Zf, c.Dxr, c.Ufr, Uf, o.H_moi[pos], FS[pos]);
                o.dH[pos] += dT / (1 - c.Dorong) * (S + tH);
            }
        }
}

/* Q-boundary inflow volume of one half-step: discharge Q(t) of each Q side
 * applied on segment 0 (inflow positive: the up/right sides face +j/+i). */
double boundary_inflow(const Oracle &o, double t_in) {
    const double t = (double)(float)t_in;
    const int t1 = (int)(t / 3600);
    const double t2 = (t - (3600.0f * t1)) / 3600.0f;
    double total = 0;
    const std::vector<Seg> *segs[4] = {&o.segV[o.M], &o.segV[2], &o.segH[2], &o.segH[o.N]};
    const double sign[4] = {-1.0, 1.0, 1.0, -1.0};
    for (int s = 0; s < 4; s++) {
        if (o.btype[s] != 2 || segs[s]->empty()) continue;
        const double *bc = o.bc[s].data();
        const int k0 = std::min(t1, o.total_time - 1), k1 = std::min(t1 + 1, o.total_time - 1);
        double q = bc[k0] * (1.0f - t2) + bc[k1] * t2;
        total += sign[s] * q * (0.5 * o.c.dT);
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* Half-steps: engine.cpp:66-170 and :176-288                          */
/* ------------------------------------------------------------------ */
void halfstep_uz(Oracle &o, double t) {
    const hd_options &ops = o.ops;
    std::vector<double> tmp;
    if (!ops.channel) update_boundary(o, t);
    int start = 2, end = o.M;
    if (ops.channel && ops.kenhhepd) {
        start = 3;
        end = o.M - 1;
    }
    uz_sweep(o, start, end);
    normalize(o, true, tmp);
    solve_v(o, 2, o.N);
    margin_v(o, 2, o.N);
    normalize(o, false, tmp);
    update_h_moi(o);
    update_uvz(o);
    find_limits(o);
    htuongdoi(o);
    o.inflow += boundary_inflow(o, t);
    if (o.cfg.sediment && t >= ops.sediment_start) {
        find_vth(o);
        heso_k(o);
        fs_sweep_j(o, t, 3, o.M - 1);
        update_fs(o);
    }
}

void halfstep_vz(Oracle &o, double t) {
    const hd_options &ops = o.ops;
    std::vector<double> tmp;
    if (!ops.channel) update_boundary(o, t);
    int start = 2, end = o.N;
    if (ops.channel && ops.kenhhepng) start = 3;
    vz_sweep(o, start, end);
    normalize(o, false, tmp);
    solve_u(o, 2, o.M);
    margin_u(o, 2, o.M);
    normalize(o, true, tmp);
    update_h_moi(o);
    reset_states_vertical(o);
    update_uvz(o);
    find_limits(o);
    htuongdoi(o);
    o.inflow += boundary_inflow(o, t);
    if (o.cfg.sediment && t >= ops.sediment_start) {
        find_vth(o);
        heso_k(o);
        fs_sweep_i(o, t, 3, o.N - 1);
        update_fs(o);
        if (o.cfg.bed_change && t >= ops.bed_change_start) {
            calculate_qb(o);
            bed_load(o);
        }
    }
}

bool copy_series(std::vector<double> &dst, const double *src, int n) {
    if (!src) {
        dst.assign(n, 0.0);
        return false;
    }
    dst.assign(src, src + n);
    return true;
}

}  // namespace

/* ================================================================== */
/* C API                                                               */
/* ================================================================== */
struct ho_ctx {
    Oracle o;
};

extern "C" {

int ho_finalize_coeffs(hd_coeffs *ce) {
    if (!ce) return HD_E_INVALID;
    const double PI = 3.14159265358979323846;
    ce->dXbp = ce->dX * ce->dX;
    ce->dX2 = 2 * ce->dX;
    ce->dYbp = ce->dY * ce->dY;
    ce->dY2 = 2 * ce->dY;
    ce->dTchia2dX = ce->dT / ce->dX2;
    ce->dTchia2dY = ce->dT / ce->dY2;
    ce->QuyDoiTime = 1.0 / 3600;
    ce->QuyDoiPi = 1.0 / PI;
    ce->HaiChiadT = 2.0 / ce->dT;
    ce->Ks = 2.5 * ce->dm;
    ce->Windx = 0.0013 * (0.00075 + 0.000067 * std::fabs(ce->Wind)) * std::fabs(ce->Wind) * ce->Wind *
                std::cos(ce->huonggio * (PI / 180));
    ce->Windy = 0.0013 * (0.00075 + 0.000067 * std::fabs(ce->Wind)) * std::fabs(ce->Wind) * ce->Wind *
                std::sin(ce->huonggio * (PI / 180));
    ce->Ufr = 0.25 * std::pow((ce->Sx - 1) * ce->g, 8.0 / 15.0) * std::pow(ce->dm, 9.0 / 15.0) *
              std::pow(ce->muy, -1.0 / 15.0);
    // Keep (Q15): the precedence of "/ muy * muy" is the reference's.
    ce->Dxr = ce->dm * std::pow(ce->g * (ce->Sx - 1) / ce->muy * ce->muy, (1.0 / 3.0));
    ce->wss = 10 * ce->muy / ce->dm *
              (std::sqrt(1 + 0.01 * (ce->Sx - 1) * 9.81 * std::pow(ce->dm, 3) / std::pow(ce->muy, 2)) - 1);
    return HD_OK;
}

int ho_load_coeffs(hd_coeffs *ce) {
    if (!ce) return HD_E_INVALID;
    std::memset(ce, 0, sizeof(*ce));
    ce->dX = 10.0;
    ce->dY = 10.0;
    ce->dT = 2;
    ce->kenhhepng = 0;
    ce->kenhhepd = 0;
    ce->NANGDAY = 0.0;
    ce->H_TINH = 0.02;
    ce->Wind = 0.0;
    ce->huonggio = 0.0;
    ce->Zbandau = 0.0;
    ce->heso = 0.94;
    ce->mu_mn = 0.2;
    ce->NDnen = 0.00;
    ce->NDbphai = 0.5;
    ce->NDbtrai = 0.5;
    ce->NDbtren = 0.5;
    ce->NDbduoi = 0.5;
    ce->Tod = 0.06;
    ce->Toe = 0.15;
    ce->hstoe = 0;
    ce->ghtoe = 3;
    ce->Mbochat = 0.00001;
    ce->ro = 1000;
    ce->ros = 2000;
    ce->dm = 0.00001;
    ce->d90 = 0.002;
    ce->muy = 1.01e-06;
    ce->Dorong = 0.5;
    ce->Sx = 2;
    ce->sohn = 8;
    ce->soha = 3;
    ce->sohn1 = 3;
    ce->CORIOLIS_FORCE = 0.0;
    ce->g = 9.81;
    return ho_finalize_coeffs(ce);
}

int ho_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return HD_OK;
#else
    (void)n;
    return HD_OK;
#endif
}

int ho_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int ho_create(const hd_mesh *mesh, const hd_coeffs *coeffs, const hd_options *ops,
              const hd_config *cfg, ho_ctx **out) {
    if (!mesh || !coeffs || !ops || !out) return HD_E_INVALID;
    if (mesh->M < 4 || mesh->N < 4 || !mesh->h || !mesh->hsnham || !mesh->VISCOIDX) return HD_E_INVALID;
    if (mesh->total_time < 2) return HD_E_INVALID;
    ho_ctx *ctx = new ho_ctx();
    Oracle &o = ctx->o;
    o.M = mesh->M;
    o.N = mesh->N;
    o.W = o.M + 3;
    o.H = o.N + 3;
    // one spare row: Onetime_init writes row N+3 (Q3)
    o.n = (size_t)(o.H + 1) * o.W;
    const size_t nin = (size_t)o.H * o.W;
    o.c = *coeffs;
    o.ops = *ops;
    o.ops.M = o.M;
    o.ops.N = o.N;
    if (cfg) o.cfg = *cfg;
    auto load = [&](std::vector<double> &dst, const double *src) {
        dst.assign(o.n, 0.0);
        if (src) std::copy(src, src + nin, dst.begin());
    };
    load(o.h, mesh->h);
    load(o.hsnham, mesh->hsnham);
    load(o.VISCOIDX, mesh->VISCOIDX);
    load(o.Fw, mesh->Fw);
    load(o.Tsxw, mesh->Tsxw);
    load(o.Tsyw, mesh->Tsyw);
    // Define (Q1): all state starts at zero unless given.
    load(o.u, mesh->u0);
    load(o.v, mesh->v0);
    load(o.z, mesh->z0);
    load(o.FS, mesh->FS0);
    for (auto *v : {&o.t_u, &o.t_v, &o.t_z, &o.H_moi, &o.Htdu, &o.Htdv, &o.tFS, &o.VTH, &o.Kx, &o.Ky,
                    &o.Qbx, &o.Qby, &o.dH, &o.htaiz, &o.htaiz_bd, &o.Kx1, &o.Ky1})
        v->assign(o.n, 0.0);
    o.khouot.assign(o.n, 0);
    o.total_time = mesh->total_time;
    for (int s = 0; s < 4; s++) {
        o.btype[s] = mesh->boundary_type[s];
        o.bienQ[s] = mesh->boundary_type[s] == 2 ? 1 : 0;
        copy_series(o.bc[s], mesh->bc[s], mesh->total_time);
        if (mesh->CC[s])
            copy_series(o.CC[s], mesh->CC[s], mesh->total_time);
        else
            o.CC[s].clear();
    }
    o.ubt.assign(o.M + 3, 0.0);
    o.ubp.assign(o.M + 3, 0.0);
    o.vbt.assign(o.N + 3, 0.0);
    o.vbd.assign(o.N + 3, 0.0);
    o.hi.assign(2 * (o.M + o.N + 6), 0.0);
    o.segH.assign(o.H, {});
    o.segV.assign(o.W, {});
    *out = ctx;
    return HD_OK;
}

int ho_init(ho_ctx *ctx) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    onetime_init(o);
    find_limits(o);
    htuongdoi(o);
    preprocess_data(o);
    o.inited = true;
    return HD_OK;
}

int ho_update_boundary(ho_ctx *ctx, double t) {
    if (!ctx) return HD_E_INVALID;
    update_boundary(ctx->o, t);
    return HD_OK;
}

int ho_halfstep_uz(ho_ctx *ctx, double t) {
    if (!ctx) return HD_E_INVALID;
    if (!ctx->o.inited) return fail(&ctx->o, HD_E_STATE, "init first");
    halfstep_uz(ctx->o, t);
    return HD_OK;
}

int ho_halfstep_vz(ho_ctx *ctx, double t) {
    if (!ctx) return HD_E_INVALID;
    if (!ctx->o.inited) return fail(&ctx->o, HD_E_STATE, "init first");
    halfstep_vz(ctx->o, t);
    return HD_OK;
}

int ho_sediment_j(ho_ctx *ctx, double t) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    find_vth(o);
    heso_k(o);
    fs_sweep_j(o, t, 3, o.M - 1);
    update_fs(o);
    return HD_OK;
}

int ho_sediment_i(ho_ctx *ctx, double t) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    find_vth(o);
    heso_k(o);
    fs_sweep_i(o, t, 3, o.N - 1);
    update_fs(o);
    return HD_OK;
}

int ho_bedload(ho_ctx *ctx, double t) {
    (void)t;
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    calculate_qb(o);
    bed_load(o);
    return HD_OK;
}

int ho_run_stage(ho_ctx *ctx, const char *name) {
    if (!ctx || !name) return HD_E_INVALID;
    Oracle &o = ctx->o;
    std::vector<double> tmp;
    const std::string n = name;
    if (n == "normalize_u") normalize(o, true, tmp);
    else if (n == "normalize_v") normalize(o, false, tmp);
    else if (n == "find_vth") find_vth(o);
    else if (n == "heso_k") heso_k(o);
    else if (n == "calculate_qb") calculate_qb(o);
    else if (n == "bed_load") bed_load(o);
    else if (n == "uz_sweep") uz_sweep(o, 2, o.M);
    else if (n == "vz_sweep") vz_sweep(o, 2, o.N);
    else if (n == "reset_states_vertical") reset_states_vertical(o);
    else if (n == "find_limits") find_limits(o);
    else return HD_E_INVALID;
    return HD_OK;
}

int ho_step(ho_ctx *ctx, double *t) {
    if (!ctx || !t) return HD_E_INVALID;
    if (!ctx->o.inited) return fail(&ctx->o, HD_E_STATE, "init first");
    const double dT = ctx->o.c.dT;
    *t += 0.5 * dT;
    halfstep_uz(ctx->o, *t);
    *t += dT * 0.5;
    halfstep_vz(ctx->o, *t);
    return HD_OK;
}

int ho_run(ho_ctx *ctx, double *t, double Tmax) {
    if (!ctx || !t) return HD_E_INVALID;
    const int T = (int)Tmax;
    while (*t < T) {
        int rc = ho_step(ctx, t);
        if (rc) return rc;
    }
    return HD_OK;
}

int ho_copy_out(ho_ctx *ctx, int field, double *dst, size_t n) {
    if (!ctx || !dst) return HD_E_INVALID;
    std::vector<double> *v = field_ptr(&ctx->o, field);
    const size_t nin = (size_t)ctx->o.H * ctx->o.W;
    if (!v || n != nin) return HD_E_INVALID;
    std::copy(v->begin(), v->begin() + nin, dst);
    return HD_OK;
}

int ho_copy_in(ho_ctx *ctx, int field, const double *src, size_t n) {
    if (!ctx || !src) return HD_E_INVALID;
    std::vector<double> *v = field_ptr(&ctx->o, field);
    const size_t nin = (size_t)ctx->o.H * ctx->o.W;
    if (!v || n != nin) return HD_E_INVALID;
    std::copy(src, src + nin, v->begin());
    return HD_OK;
}

int ho_copy_out_i32(ho_ctx *ctx, int field, int32_t *dst, size_t n) {
    if (!ctx || !dst || field != HD_F_KHOUOT) return HD_E_INVALID;
    const size_t nin = (size_t)ctx->o.H * ctx->o.W;
    if (n != nin) return HD_E_INVALID;
    std::copy(ctx->o.khouot.begin(), ctx->o.khouot.begin() + nin, dst);
    return HD_OK;
}

int ho_copy_in_i32(ho_ctx *ctx, int field, const int32_t *src, size_t n) {
    if (!ctx || !src || field != HD_F_KHOUOT) return HD_E_INVALID;
    const size_t nin = (size_t)ctx->o.H * ctx->o.W;
    if (n != nin) return HD_E_INVALID;
    std::copy(src, src + nin, ctx->o.khouot.begin());
    return HD_OK;
}

static std::vector<double> *boundary_ptr(Oracle &o, int which) {
    switch (which) {
        case HD_B_UBT: return &o.ubt;
        case HD_B_UBP: return &o.ubp;
        case HD_B_VBT: return &o.vbt;
        case HD_B_VBD: return &o.vbd;
        case HD_B_HI: return &o.hi;
        default: return nullptr;
    }
}

int ho_copy_out_boundary(ho_ctx *ctx, int which, double *dst, size_t n) {
    if (!ctx || !dst) return HD_E_INVALID;
    std::vector<double> *v = boundary_ptr(ctx->o, which);
    if (!v || n != v->size()) return HD_E_INVALID;
    std::copy(v->begin(), v->end(), dst);
    return HD_OK;
}

int ho_copy_in_boundary(ho_ctx *ctx, int which, const double *src, size_t n) {
    if (!ctx || !src) return HD_E_INVALID;
    std::vector<double> *v = boundary_ptr(ctx->o, which);
    if (!v || n != v->size()) return HD_E_INVALID;
    std::copy(src, src + n, v->begin());
    return HD_OK;
}

int ho_get_hmax(ho_ctx *ctx, double out[4]) {
    if (!ctx || !out) return HD_E_INVALID;
    for (int k = 0; k < 4; k++) out[k] = ctx->o.hmax[k];
    return HD_OK;
}

int ho_set_hmax(ho_ctx *ctx, const double in[4]) {
    if (!ctx || !in) return HD_E_INVALID;
    for (int k = 0; k < 4; k++) ctx->o.hmax[k] = in[k];
    return HD_OK;
}

int ho_get_segments(ho_ctx *ctx, int dir, int32_t *counts, size_t ncounts, int32_t *first,
                    int32_t *last, size_t cap, int64_t *total) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    const std::vector<std::vector<Seg>> &S = dir == 0 ? o.segH : o.segV;
    if (counts && ncounts != S.size()) return HD_E_INVALID;
    int64_t k = 0;
    for (size_t line = 0; line < S.size(); line++) {
        if (counts) counts[line] = (int32_t)S[line].size();
        for (const Seg &s : S[line]) {
            if ((size_t)k < cap && first && last) {
                first[k] = s.first;
                last[k] = s.last;
            }
            k++;
        }
    }
    if (total) *total = k;
    return HD_OK;
}

int ho_mass_balance(ho_ctx *ctx, double *volume, double *inflow) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    // per-row sums in column order, then the rows in order (the order the
    // library's device reduction uses, so equal states give equal volumes)
    double vol = 0;
    for (int i = 2; i <= o.N; i++) {
        double row = 0;
        for (int j = 2; j <= o.M; j++) {
            const int p = i * o.W + j;
            if (o.khouot[p] == 0) row += (o.htaiz[p] + o.z[p]) * o.c.dX * o.c.dY;
        }
        vol += row;
    }
    if (volume) *volume = vol;
    if (inflow) *inflow = o.inflow;
    return HD_OK;
}

int ho_destroy(ho_ctx *ctx) {
    delete ctx;
    return HD_OK;
}

const char *ho_last_error(ho_ctx *ctx) { return ctx ? ctx->o.err.c_str() : ""; }

int ho_thomas(int sn, const double *AA, const double *BB, const double *CC, const double *DD, double *x,
              double *Ap, double *Bp) {
    if (sn < 0 || !AA || !BB || !CC || !DD || !x || !Ap || !Bp) return HD_E_INVALID;
    thomas(sn, AA, BB, CC, DD, x, Ap, Bp);
    return HD_OK;
}

double ho_source_chs(const hd_coeffs *c, double wsm, double Zf, double fs, double fw, double vth,
                     double h_moi) {
    return source_chs(*c, wsm, Zf, fs, fw, vth, h_moi);
}

double ho_source_nchs(const hd_coeffs *c, double wsm, double Zf, double Uf, double h_moi, double fs) {
    return source_nchs(*c, wsm, Zf, c->Dxr, c->Ufr, Uf, h_moi, fs);
}

int ho_segment_line(const int32_t *line, int end, int32_t *first, int32_t *last, int cap) {
    if (!line || end < 2) return HD_E_INVALID;
    std::vector<Seg> segs;
    scan_line([&](int k) { return (int)line[k]; }, end, segs);
    for (int k = 0; k < (int)segs.size() && k < cap; k++) {
        if (first) first[k] = segs[k].first;
        if (last) last[k] = segs[k].last;
    }
    return (int)segs.size();
}

int ho_line_coeffs(ho_ctx *ctx, int dir, int line, int seg, double *out, int cap) {
    if (!ctx || !out) return HD_E_INVALID;
    Oracle &o = ctx->o;
    const auto &S = dir == 1 ? o.segV : o.segH;
    if (line < 0 || line >= (int)S.size() || seg < 0 || seg >= (int)S[line].size()) return HD_E_INVALID;
    const int len = dir == 1 ? o.N : o.M;
    if (cap < len + 4) return HD_E_INVALID;
    LineWork w;
    w.resize(len);
    const Seg s = S[line][seg];
    for (int p = s.first; p <= s.last; p++) {
        if (dir == 1)
            uz_preindex(o, p, line, w);
        else
            vz_preindex(o, line, p, w);
    }
    const std::vector<double> *arr[7] = {&w.f1, &w.f2, &w.f3, &w.f5, &w.a2, &w.c2, &w.d2};
    for (int a = 0; a < 7; a++)
        for (int p = 0; p < len + 4; p++) out[(size_t)a * cap + p] = (*arr[a])[p];
    return s.last;
}

int ho_fs_line_system(ho_ctx *ctx, int dir, int line, int seg, double t, double s_start, double *AA,
                      double *BB, double *CC, double *DD, int cap) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    const auto &S = dir == 1 ? o.segV : o.segH;
    if (line < 0 || line >= (int)S.size() || seg < 0 || seg >= (int)S[line].size()) return HD_E_INVALID;
    LineWork w;
    w.resize(dir == 1 ? o.N : o.M);
    int sn = -1;
    if (dir == 1)
        fs_line_j(o, line, t, s_start, w, seg, &sn);
    else
        fs_line_i(o, line, t, s_start, w, seg, &sn);
    if (sn + 1 > cap) return HD_E_INVALID;
    for (int r = 0; r <= sn; r++) {
        AA[r] = w.AA[r];
        BB[r] = w.BB[r];
        CC[r] = w.CC[r];
        DD[r] = w.DD[r];
    }
    return sn;
}

int ho_line_system(ho_ctx *ctx, int dir, int line, int seg, double *AA, double *BB, double *CC,
                   double *DD, double *x, int cap) {
    if (!ctx) return HD_E_INVALID;
    Oracle &o = ctx->o;
    const auto &S = dir == 1 ? o.segV : o.segH;
    if (line < 0 || line >= (int)S.size() || seg < 0 || seg >= (int)S[line].size()) return HD_E_INVALID;
    LineWork w;
    w.resize(dir == 1 ? o.N : o.M);
    int sn = -1;
    sweep_line(o, dir, line, w, seg, &sn);
    if (sn + 1 > cap) return HD_E_INVALID;
    for (int r = 0; r <= sn; r++) {
        AA[r] = w.AA[r];
        BB[r] = w.BB[r];
        CC[r] = w.CC[r];
        DD[r] = w.DD[r];
        x[r] = w.x[r];
    }
    return sn;
}

}  // extern "C"
