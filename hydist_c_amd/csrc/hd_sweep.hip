// hd_sweep.hip - implicit line sweeps of the ADI timestep on gfx950.
//
// A sweep is split into three kernels:
//   assemble: one thread per cell forms the rows of every wet segment's
//             tridiagonal system (UVZSolver_multithread.cpp:110-470 for the
//             hydraulics, sediment_transport.cpp:51-313 for FS);
//   solve:    either Thomas in the reference's row order, one lane per line
//             (bitwise FP64, UVZSolver_multithread.cpp:19-43), or, for the
//             fast forms, one workgroup per line: the partition solver
//             (k_part, default) or cyclic reduction (k_cr_reg);
//   extract:  one thread per cell writes the solution back into t_z and
//             t_u / t_v (UVZSolver_multithread.cpp:472-603) or tFS
//             (sediment_transport.cpp:167-201, :316-349).
// In the fast forms with the partition solver, a sweep whose lines are
// rows (VZ, FS along i) is one kernel: k_part assembles the rows itself
// and extracts from its LDS copy of the solution, so neither the rows nor
// the solution go through HBM.
//
// Row space of a line: position p in [2, L] (L = N along columns, M along
// rows) owns row rho = 2(p-2) for z_p and rho = 2(p-2)+1 for the velocity
// at face p (hydraulics), or rho = p-2 (FS).  A segment's system is the
// contiguous row range [rho0, rho0 + sn]; every other row is an identity
// row, so a line is one block-diagonal tridiagonal system.  The first row
// of a system has A = 0 and the last C = 0 (the reference never reads
// those entries).
//
// Layouts: ROW_MAJOR keeps row rho of all lines adjacent ([rho][line]) so
// the lanes of a wavefront (one line each) read coalesced; LINE_MAJOR keeps
// each line contiguous ([line][rho]) for the workgroup-per-line solver.
#include "hd_internal.h"

namespace hd {

// ---------------------------------------------------------------------------
// Segment description shared by assembly and extraction
// ---------------------------------------------------------------------------
// The end values a segment's boundary rows read (the water levels at its
// first and last cells, the discharge boundary values of its line) are
// loaded where a row needs them (tz_first / tz_last / ub_start / ub_end),
// not for every cell.
struct SegCtx {
    SegInfo g;
    int rho0;  // first row of the system (-1 when sn < 0)
};

template <int DIR>
HD_DEV double tz_at(const DevState &s, int line, int p) {
    return s.t_z[DIR == 1 ? (size_t)p * s.P + line : (size_t)line * s.P + p];
}
template <int DIR>
HD_DEV double ub_start(const DevState &s, int line) { return DIR == 1 ? s.ubt[line] : s.vbd[line]; }
template <int DIR>
HD_DEV double ub_end(const DevState &s, int line) { return DIR == 1 ? s.ubp[line] : s.vbt[line]; }

// Segment [f, l] of line `line` containing position p, or false; k is the
// segment's index in its line's table (columns: read only when the caller
// asks, with want_k).
template <int DIR>
HD_DEV bool locate(const DevState &s, int line, int p, int &k, int &f, int &l, bool want_k = true) {
    if (DIR == 1) {
        const size_t c = (size_t)p * s.P + line;
        const int2 fl = *reinterpret_cast<const int2 *>(s.segflV + 2 * c);
        f = fl.x;
        l = fl.y;
        k = -1;
        if (f < 0) return false;
        if (want_k) k = s.segidV[c];
        return true;
    }
    k = s.segidH[(size_t)line * s.P + p];
    if (k < 0) return false;
    const size_t o = (size_t)line * s.capH + k;
    f = s.segH_f[o];
    l = s.segH_l[o];
    return true;
}

template <int DIR>
HD_DEV SegCtx seg_ctx(const DevState &s, int line, int f, int l) {
    SegCtx c;
    const bool dkQ1 = DIR == 1 ? s.bienQ[2] : s.bienQ[1];
    const bool dkQ2 = DIR == 1 ? s.bienQ[3] : s.bienQ[0];
    const bool q11 = (s.quirks & HD_QUIRK_Q11_LITERAL) != 0;
    c.g = seg_info<DIR>(s, line, f, l, dkQ1, dkQ2, q11);
    c.rho0 = c.g.sn < 0 ? -1 : 2 * (f - 2) + (c.g.isB ? 0 : 1);
    return c;
}

// ---------------------------------------------------------------------------
// Hydraulic assembly: rows 2(p-2) (z_p) and 2(p-2)+1 (velocity at p).
// ---------------------------------------------------------------------------
struct Row {
    double a, b, c, d;
};

HD_DEV Row identity_row() { return Row{0.0, 1.0, 0.0, 0.0}; }

// Place of local row `loc` in a system whose last local row is sn (the
// plain solver's row codes): 0 outside, 1 first of several, 2 interior,
// 3 last of several, 4 the only row.
HD_DEV signed char row_code(int loc, int sn) {
    if (sn < 0 || loc < 0 || loc > sn) return 0;
    if (sn == 0) return 4;
    return loc == 0 ? 1 : (loc == sn ? 3 : 2);
}

// The two rows owned by position p of its segment [f, l] (see the layout
// rules in the reference's bienrandau / bienlongdau and end overrides); sc
// receives the segment's context.  The cell's own coefficients are formed
// first: their loads do not depend on the segment, so they are in flight
// while the segment's context is read.  FAST: the fast forms' arithmetic
// (preindex_fast, calc_abcd_fast); otherwise the reference's, bit for bit.
template <int DIR, bool FAST = false>
HD_DEV void hyd_rows(const DevState &s, int line, int p, int f, int l, double f4, Row &rz, Row &ru, SegCtx &sc) {
    const Pre cur = preindex<DIR, FAST>(s, line, p);
    const Pre nx = continuity<DIR>(s, line, p + 1);  // p + 1 <= L + 1: inside the field
    sc = seg_ctx<DIR>(s, line, f, l);
    const SegInfo &g = sc.g;
    const int sn = g.sn;
    rz = identity_row();
    ru = identity_row();
    if (sn < 0) return;
    // local row indices (relative to the system's first row)
    const int base = g.isB ? 2 * f : 2 * f + 1;  // global 2p maps to local 2p - base
    const int rz_loc = 2 * p - base, ru_loc = 2 * p + 1 - base;
    const bool hz = rz_loc >= 0 && rz_loc <= sn;
    const bool hu = ru_loc >= 0 && ru_loc <= sn;
    const Pre &nxt = p + 1 <= l ? nx : cur;
    if (hz) {
        // continuity of cell p
        Row r{cur.a2, 1.0, cur.c2, cur.d2};
        if (rz_loc == 0 && g.isB && !g.ran1) r.d = cur.d2 - cur.a2 * ub_start<DIR>(s, line);
        if (rz_loc == sn) {
            if (g.mat_end == END_Q) r.d = cur.d2 - cur.c2 * ub_end<DIR>(s, line);
            else r.d = cur.d2;  // END_WALL (END_LONG never ends on a continuity row)
        }
        if (rz_loc == 0) r.a = 0.0;
        if (rz_loc == sn) r.c = 0.0;
        rz = r;
    }
    if (hu) {
        const Mom m = FAST ? calc_abcd_fast(p, f, l, f4, g.ran1, g.ran2, cur, nxt)
                           : calc_abcd(p, f, l, f4, g.ran1, g.ran2, cur, nxt);
        Row r{m.a1, m.b1, m.c1, m.d1};
        if (ru_loc == 0 && !g.isB) r.d = m.d1 - m.a1 * tz_at<DIR>(s, line, f);
        if (ru_loc == sn) r.d = m.d1 - m.c1 * tz_at<DIR>(s, line, l);  // END_LONG
        if (ru_loc == 0) r.a = 0.0;
        if (ru_loc == sn) r.c = 0.0;
        ru = r;
    }
}

// Row-major index of row rho of line `line`; line-major index.
HD_DEV size_t rm_idx(int rho, int line, int ld) { return (size_t)rho * ld + line; }
HD_DEV size_t lm_idx(int rho, int line, int ld) { return (size_t)line * ld + rho; }

// Threads: x along j (coalesced state reads), y along i.  For outputs whose
// layout runs against the thread order, a 64 x 16 tile is transposed
// through LDS.
template <int DIR, int LAYOUT>
__global__ void __launch_bounds__(512) k_hyd_assemble(const DevState *__restrict__ S, SweepBufs Bin, int start,
                                                      int end) {
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 8 + threadIdx.y;
    const int line = DIR == 1 ? j : i;
    const int p = DIR == 1 ? i : j;
    const int L = DIR == 1 ? s.N : s.M;
    const bool active = line >= start && line <= end && p >= 2 && p <= L;
    const double f4 = DIR == 1 ? 2 * s.c.g * s.c.dTchia2dX : 2 * s.c.g * s.c.dTchia2dY;
    Row rz = identity_row(), ru = identity_row();
    signed char cz = 0, cu = 0;
    if (active) {
        int k, f, l;
        if (locate<DIR>(s, line, p, k, f, l)) {
            SegCtx sc;
            hyd_rows<DIR, LAYOUT == LINE_MAJOR>(s, line, p, f, l, f4, rz, ru, sc);
            const int base = sc.g.isB ? 2 * f : 2 * f + 1;  // as in hyd_rows
            cz = row_code(2 * p - base, sc.g.sn);
            cu = row_code(2 * p + 1 - base, sc.g.sn);
            if (p == f) {
                const int cap = DIR == 1 ? s.capV : s.capH;
                B.sys0[(size_t)line * cap + k] = sc.rho0;
                B.sysn[(size_t)line * cap + k] = sc.g.sn;
            }
        }
    }
    // Output orientation: the thread grid runs along j.  UZ lines are
    // columns (line = j): ROW_MAJOR [rho][line] is coalesced directly.
    // VZ lines are rows (line = i, rho from j): LINE_MAJOR [line][rho] is
    // coalesced directly.  The other two combinations go through LDS.
    constexpr bool direct = (DIR == 1 && LAYOUT == ROW_MAJOR) || (DIR == 0 && LAYOUT == LINE_MAJOR);
    if (direct) {
        if (!active) return;
        const int rz_i = 2 * (p - 2), ru_i = rz_i + 1;
        const int ld = B.ld;
        size_t iz, iu;
        if (LAYOUT == ROW_MAJOR) {
            iz = rm_idx(rz_i, line, ld);
            iu = rm_idx(ru_i, line, ld);
        } else {
            iz = lm_idx(rz_i, line, ld);
            iu = lm_idx(ru_i, line, ld);
        }
        if (LAYOUT == LINE_MAJOR) {
            // rows 2(p-2) and 2(p-2)+1 are adjacent: one 16-byte store per
            // array (iz is even and ld, the scenario stride and the buffers
            // are multiples of 64 elements, so the pair is 16-byte aligned)
            *reinterpret_cast<double2 *>(B.A + iz) = make_double2(rz.a, ru.a);
            *reinterpret_cast<double2 *>(B.B + iz) = make_double2(rz.b, ru.b);
            *reinterpret_cast<double2 *>(B.C + iz) = make_double2(rz.c, ru.c);
            *reinterpret_cast<double2 *>(B.D + iz) = make_double2(rz.d, ru.d);
        } else {
            B.A[iz] = rz.a; B.B[iz] = rz.b; B.C[iz] = rz.c; B.D[iz] = rz.d;
            B.A[iu] = ru.a; B.B[iu] = ru.b; B.C[iu] = ru.c; B.D[iu] = ru.d;
            B.code[iz] = cz;
            B.code[iu] = cu;
        }
    } else {
        // tile: 8 (i) x 64 (j) cells, 2 rows each -> transpose so that the
        // runs along i become contiguous
        __shared__ double tA[8][129], tB[8][129], tC[8][129], tD[8][129];
        __shared__ signed char tK[8][128];
        const int ty = threadIdx.y, tx = threadIdx.x;
        tA[ty][2 * tx] = rz.a; tB[ty][2 * tx] = rz.b; tC[ty][2 * tx] = rz.c; tD[ty][2 * tx] = rz.d;
        tA[ty][2 * tx + 1] = ru.a; tB[ty][2 * tx + 1] = ru.b; tC[ty][2 * tx + 1] = ru.c; tD[ty][2 * tx + 1] = ru.d;
        tK[ty][2 * tx] = cz;
        tK[ty][2 * tx + 1] = cu;
        __syncthreads();
        // write out: for DIR==1 & LINE_MAJOR: line = j (64 lines), rho from i (8 cells -> 16 rows)
        //            for DIR==0 & ROW_MAJOR:  line = i (8 lines), rho from j (64 cells -> 128 rows)
        const int t = ty * 64 + tx;  // 0..511
        const int i0 = blockIdx.y * 8, j0 = blockIdx.x * 64;
        const int ld = B.ld;
        if (DIR == 1) {
            // 64 lines x 16 rows = 1024 values per array, 2 per thread
            for (int e = t; e < 64 * 16; e += 512) {
                const int ln = e / 16, r = e % 16;  // r runs fastest -> contiguous in [line][rho]
                const int jj = j0 + ln, ci = i0 + r / 2, slot = r & 1;
                if (jj < start || jj > end || ci < 2 || ci > L) continue;
                const size_t o = lm_idx(2 * (ci - 2) + slot, jj, ld);
                const int col = 2 * ln + slot;
                B.A[o] = tA[r / 2][col]; B.B[o] = tB[r / 2][col]; B.C[o] = tC[r / 2][col]; B.D[o] = tD[r / 2][col];
            }
        } else {
            // 8 lines x 128 rows; line index fastest for coalescing in [rho][line]
            for (int e = t; e < 8 * 128; e += 512) {
                const int ln = e % 8, r = e / 8;
                const int ii = i0 + ln, cj = j0 + r / 2, slot = r & 1;
                if (ii < start || ii > end || cj < 2 || cj > L) continue;
                const size_t o = rm_idx(2 * (cj - 2) + slot, ii, ld);
                B.A[o] = tA[ln][r]; B.B[o] = tB[ln][r]; B.C[o] = tC[ln][r]; B.D[o] = tD[ln][r];
                B.code[o] = tK[ln][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FS assembly: row p-2 for cells in [f+1, l-1] of segments of >= 3 cells.
// ---------------------------------------------------------------------------
template <int DIR, int LAYOUT>
__global__ void __launch_bounds__(512) k_fs_assemble(const DevState *__restrict__ S, SweepBufs Bin, int start,
                                                     int end, double t_in, double s_start, int from_clock) {
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    const double t = from_clock ? s.clock[0] : t_in;
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 8 + threadIdx.y;
    const int line = DIR == 1 ? j : i;
    const int p = DIR == 1 ? i : j;
    const int L = DIR == 1 ? s.N : s.M;
    const bool active = line >= start && line <= end && p >= 2 && p <= L;
    Row r = identity_row();
    signed char code = 0;
    if (active) {
        int k, f, l;
        if (locate<DIR>(s, line, p, k, f, l) && l - f >= 2) {
            bool ran1, ran2;
            wall_flags(s, DIR, line, f, l, ran1, ran2);
            code = row_code(p - (f + 1), l - f - 2);
            if (p >= f + 1 && p <= l - 1) {
                if (DIR == 1)
                    fs_row<LAYOUT == LINE_MAJOR>(1, s, p, line, f, l, ran1, ran2, t, s_start, r.a, r.b, r.c, r.d);
                else
                    fs_row<LAYOUT == LINE_MAJOR>(0, s, line, p, f, l, ran1, ran2, t, s_start, r.a, r.b, r.c, r.d);
                if (p == f + 1) r.a = 0.0;
                if (p == l - 1) r.c = 0.0;
            }
            if (p == f) {
                const int cap = DIR == 1 ? s.capV : s.capH;
                B.sys0[(size_t)line * cap + k] = f + 1 - 2;
                B.sysn[(size_t)line * cap + k] = l - f - 2;
            }
        } else if (k >= 0 && p == f) {
            const int cap = DIR == 1 ? s.capV : s.capH;
            B.sys0[(size_t)line * cap + k] = -1;
            B.sysn[(size_t)line * cap + k] = -1;
        }
    }
    constexpr bool direct = (DIR == 1 && LAYOUT == ROW_MAJOR) || (DIR == 0 && LAYOUT == LINE_MAJOR);
    if (direct) {
        if (!active) return;
        const size_t o = LAYOUT == ROW_MAJOR ? rm_idx(p - 2, line, B.ld) : lm_idx(p - 2, line, B.ld);
        B.A[o] = r.a; B.B[o] = r.b; B.C[o] = r.c; B.D[o] = r.d;
        if (LAYOUT == ROW_MAJOR) B.code[o] = code;
    } else {
        __shared__ double tA[8][65], tB[8][65], tC[8][65], tD[8][65];
        __shared__ signed char tK[8][64];
        const int ty = threadIdx.y, tx = threadIdx.x;
        tA[ty][tx] = r.a; tB[ty][tx] = r.b; tC[ty][tx] = r.c; tD[ty][tx] = r.d;
        tK[ty][tx] = code;
        __syncthreads();
        const int t = ty * 64 + tx;
        const int i0 = blockIdx.y * 8, j0 = blockIdx.x * 64;
        const int ln = DIR == 1 ? t / 8 : t % 8;
        const int rr = DIR == 1 ? t % 8 : t / 8;
        if (DIR == 1) {  // line = j (64), rho from i (8), LINE_MAJOR
            const int jj = j0 + ln, ci = i0 + rr;
            if (jj < start || jj > end || ci < 2 || ci > L) return;
            const size_t o = lm_idx(ci - 2, jj, B.ld);
            B.A[o] = tA[rr][ln]; B.B[o] = tB[rr][ln]; B.C[o] = tC[rr][ln]; B.D[o] = tD[rr][ln];
        } else {  // line = i (8), rho from j (64), ROW_MAJOR
            const int ii = i0 + ln, cj = j0 + rr;
            if (ii < start || ii > end || cj < 2 || cj > L) return;
            const size_t o = rm_idx(cj - 2, ii, B.ld);
            B.A[o] = tA[ln][rr]; B.B[o] = tB[ln][rr]; B.C[o] = tC[ln][rr]; B.D[o] = tD[ln][rr];
            B.code[o] = tK[ln][rr];
        }
    }
}

// ---------------------------------------------------------------------------
// Column lines (UZ and FS along j) into [line][rho]: the state is read with
// threads along j, so each line's rows leave through an LDS transpose.
// A tile covers 64 lines and TR = 8 * CPT cells along i, starting at
// i0 = 2 + TR * blockIdx.y so that each line's run of rows starts on a
// 64 * CPT-byte (FS) or 128 * CPT-byte (hydraulics) boundary; each thread
// assembles CPT cells (i0 + threadIdx.y + 8c).  Two of the four coefficient
// arrays pass through LDS at a time, which halves the LDS of a block
// against staging all four.  Same rows as k_hyd_assemble / k_fs_assemble
// with DIR = 1.  One cell per thread measured fastest on the 2048 x 1024
// mesh (against 2 and 4: more blocks in flight beat longer write runs).
// ---------------------------------------------------------------------------
constexpr int COLS_HYD_CPT = 1;
constexpr int COLS_FS_CPT = 1;

template <int CPT>
__global__ void __launch_bounds__(512) k_hyd_assemble_cols(const DevState *__restrict__ S, SweepBufs Bin,
                                                           int start, int end) {
    constexpr int TR = 8 * CPT;
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int j = blockIdx.x * 64 + tx;
    const int i0 = 2 + blockIdx.y * TR;
    const int L = s.N;
    const double f4 = 2 * s.c.g * s.c.dTchia2dX;
    Row rz[CPT], ru[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int i = i0 + ty + 8 * c;
        rz[c] = identity_row();
        ru[c] = identity_row();
        int k, f, l;
        if (j >= start && j <= end && i <= L && locate<1>(s, j, i, k, f, l, false)) {
            SegCtx sc;
            hyd_rows<1, true>(s, j, i, f, l, f4, rz[c], ru[c], sc);
            if (i == f) {
                k = s.segidV[(size_t)i * s.P + j];
                B.sys0[(size_t)j * s.capV + k] = sc.rho0;
                B.sysn[(size_t)j * s.capV + k] = sc.g.sn;
            }
        }
    }
    __shared__ double t0[TR][2 * 64 + 1], t1[TR][2 * 64 + 1];
    const int t = ty * 64 + tx, ld = B.ld;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int cy = ty + 8 * c;
            t0[cy][2 * tx] = pass ? rz[c].c : rz[c].a;
            t0[cy][2 * tx + 1] = pass ? ru[c].c : ru[c].a;
            t1[cy][2 * tx] = pass ? rz[c].d : rz[c].b;
            t1[cy][2 * tx + 1] = pass ? ru[c].d : ru[c].b;
        }
        __syncthreads();
        double *o0 = pass ? B.C : B.A, *o1 = pass ? B.D : B.B;
        // each thread writes cell ci's two rows (z, velocity) of one line
        // as one 16-byte store per array: 2(ci-2) is even and ld, the
        // scenario stride and the buffers are multiples of 64 elements
        for (int e = t; e < 64 * TR; e += 512) {
            const int ln = e / TR, c = e % TR;  // c runs fastest: contiguous in [line][rho]
            const int jj = blockIdx.x * 64 + ln, ci = i0 + c;
            if (jj < start || jj > end || ci > L) continue;
            const size_t o = lm_idx(2 * (ci - 2), jj, ld);
            *reinterpret_cast<double2 *>(o0 + o) = make_double2(t0[c][2 * ln], t0[c][2 * ln + 1]);
            *reinterpret_cast<double2 *>(o1 + o) = make_double2(t1[c][2 * ln], t1[c][2 * ln + 1]);
        }
        if (pass == 0) __syncthreads();
    }
}

template <int CPT>
__global__ void __launch_bounds__(512) k_fs_assemble_cols(const DevState *__restrict__ S, SweepBufs Bin, int start,
                                                          int end, double t_in, double s_start, int from_clock) {
    constexpr int TR = 8 * CPT;
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    const double tm = from_clock ? s.clock[0] : t_in;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int j = blockIdx.x * 64 + tx;
    const int i0 = 2 + blockIdx.y * TR;
    const int L = s.N;
    Row rw[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int i = i0 + ty + 8 * c;
        rw[c] = identity_row();
        if (j < start || j > end || i > L) continue;
        int k, f, l;
        const bool in = locate<1>(s, j, i, k, f, l, false);
        if (in && i == f) k = s.segidV[(size_t)i * s.P + j];
        if (in && l - f >= 2) {
            bool ran1, ran2;
            wall_flags(s, 1, j, f, l, ran1, ran2);
            if (i >= f + 1 && i <= l - 1) {
                fs_row_j_fast(s, i, j, f, l, ran1, ran2, tm, s_start, rw[c].a, rw[c].b, rw[c].c, rw[c].d);
                if (i == f + 1) rw[c].a = 0.0;
                if (i == l - 1) rw[c].c = 0.0;
            }
            if (i == f) {
                B.sys0[(size_t)j * s.capV + k] = f + 1 - 2;
                B.sysn[(size_t)j * s.capV + k] = l - f - 2;
            }
        } else if (in && i == f) {
            B.sys0[(size_t)j * s.capV + k] = -1;
            B.sysn[(size_t)j * s.capV + k] = -1;
        }
    }
    __shared__ double t0[TR][64 + 1], t1[TR][64 + 1];
    const int t = ty * 64 + tx, ld = B.ld;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int cy = ty + 8 * c;
            t0[cy][tx] = pass ? rw[c].c : rw[c].a;
            t1[cy][tx] = pass ? rw[c].d : rw[c].b;
        }
        __syncthreads();
        double *o0 = pass ? B.C : B.A, *o1 = pass ? B.D : B.B;
        for (int e = t; e < 64 * TR; e += 512) {
            const int ln = e / TR, r = e % TR;
            const int jj = blockIdx.x * 64 + ln, ci = i0 + r;
            if (jj < start || jj > end || ci > L) continue;
            const size_t o = lm_idx(ci - 2, jj, ld);
            o0[o] = t0[r][ln];
            o1[o] = t1[r][ln];
        }
        if (pass == 0) __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Thomas, one lane per line, rows of all lines in lockstep ([rho][line]).
// Each row carries its place in its system (B.code, written by the
// assembly), so the recurrence needs no segment table: the first, interior
// and last rows of tridiag differ only in which operands feed its two
// divisions (selects, one division site for all lanes), and every load and
// store of a full block of TB rows is unconditional, so the next block's
// rows are in flight while the current block's serial recurrence runs.
// Lanes past the last line redo the last line (identical values to the
// same addresses).  Arithmetic: tridiag (UVZSolver_multithread.cpp:19-43),
// unchanged: den is b, ep = a*Ap + b, or b + a*Ap (the same value).
// ---------------------------------------------------------------------------
constexpr int TB = 12;  // rows per block (measured: 8 -> 12 saves 5 %, 16 no better)

struct ThomasRow {
    double a, b, c, d;
    int code;
};

__global__ void __launch_bounds__(64) k_thomas(SweepBufs Bin, int start, int end, int nrows) {
    const SweepBufs B = Bin.scen(blockIdx.y);
    const int line = min(start + (int)(blockIdx.x * 64 + threadIdx.x), end);
    const int ld = B.ld;
    const double *__restrict__ A = B.A + line;
    const double *__restrict__ Bv = B.B + line;
    const double *__restrict__ C = B.C + line;
    const double *__restrict__ D = B.D + line;
    const signed char *__restrict__ K = B.code + line;
    double *__restrict__ Ap = B.Ap + line;
    double *__restrict__ Bp = B.Bp + line;
    double *__restrict__ X = B.X + line;
    const int nfull = nrows / TB;
    auto load_row = [&](int rho) {
        const size_t o = (size_t)min(max(rho, 0), nrows - 1) * ld;
        return ThomasRow{A[o], Bv[o], C[o], D[o], (int)K[o]};
    };

    // forward
    double ApP = 0, BpP = 0;
    auto fwd = [&](const ThomasRow &w, double &ap, double &bp) {
        const bool first = w.code == 1 || w.code == 4;
        const double den = first ? w.b : w.a * ApP + w.b;
        const double num = first ? w.d : w.d - (w.a * BpP);
        ap = (w.code == 1 || w.code == 2) ? -w.c / den : 0.0;
        bp = num / den;
        if (w.code != 0) {
            ApP = ap;
            BpP = bp;
        }
    };
    ThomasRow cur[TB], nxt[TB];
#pragma unroll
    for (int q = 0; q < TB; q++) nxt[q] = load_row(q);
    for (int blk = 0; blk < nfull; blk++) {
#pragma unroll
        for (int q = 0; q < TB; q++) cur[q] = nxt[q];
#pragma unroll
        for (int q = 0; q < TB; q++) nxt[q] = load_row((blk + 1) * TB + q);
#pragma unroll
        for (int q = 0; q < TB; q++) {
            const size_t o = (size_t)(blk * TB + q) * ld;
            double ap, bp;
            fwd(cur[q], ap, bp);
            Ap[o] = ap;
            Bp[o] = bp;
        }
    }
#pragma unroll
    for (int q = 0; q < TB; q++) {  // rows past the full blocks
        const int rho = nfull * TB + q;
        if (rho >= nrows) break;
        double ap, bp;
        fwd(nxt[q], ap, bp);
        Ap[(size_t)rho * ld] = ap;
        Bp[(size_t)rho * ld] = bp;
    }

    // backward (rows outside any system get finite values nobody reads)
    double xnext = 0;
    auto bwd = [&](int code, double pa, double pb) {
        const double xv = (code == 3 || code == 4) ? pb : pb + (pa * xnext);
        xnext = xv;
        return xv;
    };
    for (int rho = nrows - 1; rho >= nfull * TB; rho--) {
        const size_t o = (size_t)rho * ld;
        X[o] = bwd((int)K[o], Ap[o], Bp[o]);
    }
    double pa[TB], pb[TB], qa[TB], qb[TB];
    int pk[TB], qk[TB];
#pragma unroll
    for (int q = 0; q < TB; q++) {
        const size_t o = (size_t)max(nfull * TB - 1 - q, 0) * ld;
        qa[q] = Ap[o];
        qb[q] = Bp[o];
        qk[q] = K[o];
    }
    for (int blk = nfull - 1; blk >= 0; blk--) {
#pragma unroll
        for (int q = 0; q < TB; q++) {
            pa[q] = qa[q];
            pb[q] = qb[q];
            pk[q] = qk[q];
        }
#pragma unroll
        for (int q = 0; q < TB; q++) {
            const size_t o = (size_t)max(blk * TB - 1 - q, 0) * ld;
            qa[q] = Ap[o];
            qb[q] = Bp[o];
            qk[q] = K[o];
        }
#pragma unroll
        for (int q = 0; q < TB; q++) X[(size_t)(blk * TB + TB - 1 - q) * ld] = bwd(pk[q], pa[q], pb[q]);
    }
}

// ---------------------------------------------------------------------------
// Workgroup-per-line solvers of the fast forms ([line][rho]).
//
// Hydraulics (INTERLEAVED): the velocity rows (odd rho) are first
// eliminated into the z rows, giving the z-only system of the line
// (strictly diagonally dominant), which is solved in precision T; with
// REFINE, the FP64 residual of that system is solved once more (reusing
// the stored elimination) and added.  The velocities are then recovered in
// FP64 from their own rows.  FS (!INTERLEAVED): the line's system is
// solved directly.  cr_solve / cr_replay below are the cyclic reduction of
// n rows held in LDS that k_cr_reg uses for the whole line; k_part solves
// its reduced system (one row per thread) by parallel cyclic reduction,
// pcr_solve / pcr_replay.
// ---------------------------------------------------------------------------
template <typename T>
struct CRShared {
    T *a, *b, *c, *d, *k1, *k2;
};

// Forward reduction + back substitution of n rows held in LDS.  When
// `store` is set, the multipliers of every update are kept in k1/k2 so a
// second right-hand side can be reduced with cr_replay.
template <typename T>
HD_DEV void cr_solve(CRShared<T> sh, int n, bool store) {
    const int tid = threadIdx.x, nt = blockDim.x;
    int K = 0;
    while ((2 << K) <= n) K++;  // 2^K <= n < 2^(K+1)
    int off = 0;
    for (int lev = 1; lev <= K; lev++) {
        const int h = 1 << (lev - 1), st = 1 << lev;
        const int cnt = n >> lev;
        for (int m = tid; m < cnt; m += nt) {
            const int i = (m + 1) * st - 1;
            const int il = i - h, ir = i + h;
            const T k1 = sh.a[i] / sh.b[il];
            T k2 = 0;
            T nb = sh.b[i] - sh.c[il] * k1;
            T nd = sh.d[i] - sh.d[il] * k1;
            T na = -sh.a[il] * k1;
            T nc = 0;
            if (ir < n) {
                k2 = sh.c[i] / sh.b[ir];
                nb = nb - sh.a[ir] * k2;
                nd = nd - sh.d[ir] * k2;
                nc = -sh.c[ir] * k2;
            }
            sh.a[i] = na;
            sh.b[i] = nb;
            sh.c[i] = nc;
            sh.d[i] = nd;
            if (store) {
                sh.k1[off + m] = k1;
                sh.k2[off + m] = k2;
            }
        }
        off += cnt;
        __syncthreads();
    }
    // top row
    if (tid == 0) {
        const int i = (1 << K) - 1;
        sh.d[i] = sh.d[i] / sh.b[i];
    }
    __syncthreads();
    for (int lev = K; lev >= 1; lev--) {
        const int h = 1 << (lev - 1), st = 1 << lev;
        // rows (i+1) % h == 0 and (i+1) % st != 0
        const int cnt = (n + h) / st;  // i = (2m+1)h - 1 < n
        for (int m = tid; m < cnt; m += nt) {
            const int i = (2 * m + 1) * h - 1;
            if (i >= n) continue;
            const int il = i - h, ir = i + h;
            T v = sh.d[i];
            if (il >= 0) v = v - sh.a[i] * sh.d[il];
            if (ir < n) v = v - sh.c[i] * sh.d[ir];
            sh.d[i] = v / sh.b[i];
        }
        __syncthreads();
    }
}

// Reduce and solve a new right-hand side in sh.d with the multipliers and
// final rows left by cr_solve(store = true).
template <typename T>
HD_DEV void cr_replay(CRShared<T> sh, int n) {
    const int tid = threadIdx.x, nt = blockDim.x;
    int K = 0;
    while ((2 << K) <= n) K++;
    int off = 0;
    for (int lev = 1; lev <= K; lev++) {
        const int h = 1 << (lev - 1), st = 1 << lev;
        const int cnt = n >> lev;
        for (int m = tid; m < cnt; m += nt) {
            const int i = (m + 1) * st - 1;
            const int il = i - h, ir = i + h;
            T nd = sh.d[i] - sh.d[il] * sh.k1[off + m];
            if (ir < n) nd = nd - sh.d[ir] * sh.k2[off + m];
            sh.d[i] = nd;
        }
        off += cnt;
        __syncthreads();
    }
    if (tid == 0) {
        const int i = (1 << K) - 1;
        sh.d[i] = sh.d[i] / sh.b[i];
    }
    __syncthreads();
    for (int lev = K; lev >= 1; lev--) {
        const int h = 1 << (lev - 1), st = 1 << lev;
        const int cnt = (n + h) / st;
        for (int m = tid; m < cnt; m += nt) {
            const int i = (2 * m + 1) * h - 1;
            if (i >= n) continue;
            const int il = i - h, ir = i + h;
            T v = sh.d[i];
            if (il >= 0) v = v - sh.a[i] * sh.d[il];
            if (ir < n) v = v - sh.c[i] * sh.d[ir];
            sh.d[i] = v / sh.b[i];
        }
        __syncthreads();
    }
}

// Parallel cyclic reduction of k_part's reduced system: n = blockDim.x
// rows, row t held by thread t.  Step s eliminates from every row its
// couplings to the rows h = 2^s away on both sides (rows beyond the ends
// act as identity rows), so after ceil(log2 n) steps every row stands
// alone: no back substitution, every thread busy at every step, and the
// reads of a step are contiguous neighbour rows (no LDS bank conflicts).
// A step reads the rows from one buffer and writes them to the other, so
// one barrier per step keeps the reads and writes apart.  The solution
// goes to x.  With `store`, each step's multipliers and the final pivots
// are kept so pcr_replay can solve a second right-hand side.
// LDS layout from `base` (n elements each): buffer u's a, b, c, d at
// 4u .. 4u + 3, then x, the final pivots and, per step, k1 and k2.  The
// buffers are addressed by offsets, not by an array of pointers, which
// would sit in scratch memory.
template <typename T>
struct PCRShared {
    T *base;
    int n;
    HD_DEV T *row(int buf, int field) const { return base + (4 * buf + field) * n; }
    HD_DEV T *x() const { return base + 8 * n; }
    HD_DEV T *bf() const { return base + 9 * n; }
    HD_DEV T *k(int step, int which) const { return base + (10 + 2 * step + which) * n; }
};

// Steps of the reduction of n rows: ceil(log2 n).  LDS elements of T:
// 10 n, plus 2 n per step when multipliers are kept.
__host__ __device__ inline int pcr_steps(int n) {
    int s = 0;
    while ((1 << s) < n) s++;
    return s;
}

// Solves the system whose row t is (a, b, c, d) of thread t; ends with a
// barrier, after which sh.x() holds the solution.
template <typename T>
HD_DEV void pcr_solve(const PCRShared<T> &sh, T a, T b, T c, T d, bool store) {
    const int i = threadIdx.x, n = sh.n;
    int cur = 0, s = 0;
    sh.row(0, 0)[i] = a;
    sh.row(0, 1)[i] = b;
    sh.row(0, 2)[i] = c;
    sh.row(0, 3)[i] = d;
    __syncthreads();
    for (int h = 1; h < n; h <<= 1, s++) {
        const bool hl = i >= h, hr = i + h < n;
        const int il = hl ? i - h : i, ir = hr ? i + h : i;
        const T *ra = sh.row(cur, 0), *rb = sh.row(cur, 1), *rc = sh.row(cur, 2), *rd = sh.row(cur, 3);
        const T al = ra[il], bl = rb[il], cl = rc[il], dl = rd[il];
        const T ar = ra[ir], br = rb[ir], cr = rc[ir], dr = rd[ir];
        const T k1 = hl ? a / bl : T(0);
        const T k2 = hr ? c / br : T(0);
        b = b - cl * k1 - ar * k2;
        d = d - dl * k1 - dr * k2;
        a = -(al * k1);
        c = -(cr * k2);
        if (store) {
            sh.k(s, 0)[i] = k1;
            sh.k(s, 1)[i] = k2;
        }
        if (2 * h >= n) break;  // the last step's rows are read by no one
        cur ^= 1;
        sh.row(cur, 0)[i] = a;
        sh.row(cur, 1)[i] = b;
        sh.row(cur, 2)[i] = c;
        sh.row(cur, 3)[i] = d;
        __syncthreads();
    }
    sh.x()[i] = d / b;
    if (store) sh.bf()[i] = b;
    __syncthreads();
}

// Solves the system of the last pcr_solve(store = true) for the right-hand
// side d of thread t; ends with a barrier, after which sh.x() holds the
// solution.  Only the d buffers, x and the caller's data change.
template <typename T>
HD_DEV void pcr_replay(const PCRShared<T> &sh, T d) {
    const int i = threadIdx.x, n = sh.n;
    int cur = 0, s = 0;
    sh.row(0, 3)[i] = d;
    __syncthreads();
    for (int h = 1; h < n; h <<= 1, s++) {
        const T *rd = sh.row(cur, 3);
        const T dl = rd[i >= h ? i - h : i], dr = rd[i + h < n ? i + h : i];
        d = d - dl * sh.k(s, 0)[i] - dr * sh.k(s, 1)[i];
        if (2 * h >= n) break;
        cur ^= 1;
        sh.row(cur, 3)[i] = d;
        __syncthreads();
    }
    sh.x()[i] = d / sh.bf()[i];
    __syncthreads();
}

// A z row of an interleaved line with its neighbouring velocity rows
// eliminated in FP64: zr is row 2q, m1 row 2q-1 (used when has_m1), p1 row
// 2q+1 (used when has_p1).
HD_DEV Row reduce_z_vals(const Row &m1, const Row &zr, const Row &p1, bool has_m1, bool has_p1) {
    Row o{0.0, zr.b, 0.0, zr.d};
    if (has_m1 && zr.a != 0.0) {
        const double al = zr.a / m1.b;
        o.a = -al * m1.a;
        o.b = o.b - al * m1.c;
        o.d = o.d - al * m1.d;
    }
    if (has_p1 && zr.c != 0.0) {
        const double ga = zr.c / p1.b;
        o.c = -ga * p1.c;
        o.b = o.b - ga * p1.a;
        o.d = o.d - ga * p1.d;
    }
    return o;
}

// z-row q of an interleaved line (rows 2q-1, 2q, 2q+1) reduced in FP64.
HD_DEV Row reduce_z_row(const double *__restrict__ A, const double *__restrict__ Bv,
                        const double *__restrict__ C, const double *__restrict__ D, int q, int nrows) {
    const int r = 2 * q;
    const bool has_m1 = r - 1 >= 0, has_p1 = r + 1 < nrows;
    const Row zr{A[r], Bv[r], C[r], D[r]};
    const Row m1 = has_m1 ? Row{A[r - 1], Bv[r - 1], C[r - 1], D[r - 1]} : identity_row();
    const Row p1 = has_p1 ? Row{A[r + 1], Bv[r + 1], C[r + 1], D[r + 1]} : identity_row();
    return reduce_z_vals(m1, zr, p1, has_m1, has_p1);
}

// Segments of the K positions p = tid + m * nt + 2 of row line `line` that
// a k_part thread assembles: k (the index in the line's table, -1 where p
// is in no segment) and the bounds f, l.  All K index loads go out at
// once, then all K table loads (indices clamped, so both rounds are
// unconditional), so the positions wait for two memory latencies together
// instead of two each.
template <int K>
HD_DEV void row_segments(const DevState &s, int line, int tid, int nt, int (&kk)[K], int (&ff)[K], int (&ll)[K]) {
#pragma unroll
    for (int m = 0; m < K; m++) kk[m] = s.segidH[(size_t)line * s.P + min(tid + m * nt + 2, s.M)];
#pragma unroll
    for (int m = 0; m < K; m++) {
        const size_t o = (size_t)line * s.capH + max(kk[m], 0);
        ff[m] = s.segH_f[o];
        ll[m] = s.segH_l[o];
    }
#pragma unroll
    for (int m = 0; m < K; m++)
        if (kk[m] < 0) ff[m] = ll[m] = 2;  // in-range bounds for the callers' unconditional loads
}

// ---------------------------------------------------------------------------
// Extraction (hydraulics).  Define: cells before `last` are written before
// the end cell, whose continuity formula uses the new velocity at last-1.
// ---------------------------------------------------------------------------
// Cell p of line `line`; xat(rho) reads the line's solution row rho.
template <int DIR, typename XAt>
HD_DEV void hyd_extract_at(const DevState &s, int line, int p, int f, int l, XAt xat);
template <int DIR, typename XAt>
HD_DEV void hyd_extract_cell(const DevState &s, int line, int p, XAt xat) {
    int k, f, l;
    if (locate<DIR>(s, line, p, k, f, l, false)) hyd_extract_at<DIR>(s, line, p, f, l, xat);
}
// The same for a cell whose segment [f, l] is known.
template <int DIR, typename XAt>
HD_DEV void hyd_extract_at(const DevState &s, int line, int p, int f, int l, XAt xat) {
    const int L = DIR == 1 ? s.N : s.M;
    const SegCtx sc = seg_ctx<DIR>(s, line, f, l);
    const SegInfo &g = sc.g;
    const int W = s.P;
    auto cell = [&](int q) { return DIR == 1 ? q * W + line : line * W + q; };
    double *__restrict__ tz = s.t_z;
    double *__restrict__ tv = DIR == 1 ? s.t_u : s.t_v;
    const int sn = g.sn;
    const int rz = 2 * (p - 2), ru = rz + 1;
    if (p < l) {
        if (g.isB) {
            tz[cell(p)] = xat(rz);
            tv[cell(p)] = xat(ru);
            if (p == f) tv[cell(f - 1)] = g.ran1 ? 0.0 : ub_start<DIR>(s, line);
        } else {
            tv[cell(p)] = xat(ru);
            if (p > f) tz[cell(p)] = xat(rz);
            if (p == f) {
                const Pre e = continuity<DIR>(s, line, f);
                tv[cell(f - 1)] = (e.d2 - tz[cell(f)] - e.c2 * xat(ru)) / e.a2;
            }
        }
    } else {  // p == l
        const bool ext_Q = (DIR == 1 ? s.bienQ[3] : s.bienQ[0]) && l == L;
        if (g.ran2) {
            tv[cell(l)] = 0;
            if (sn >= 0) tz[cell(l)] = xat(sc.rho0 + sn);
        } else if (ext_Q) {
            tv[cell(l)] = ub_end<DIR>(s, line);
            if (sn >= 0) tz[cell(l)] = xat(sc.rho0 + sn);
        } else {
            const Pre e = continuity<DIR>(s, line, l);
            const double prev = l > f ? xat(2 * (l - 1 - 2) + 1) : tv[cell(l - 1)];
            tv[cell(l)] = (e.d2 - e.a2 * prev - tz[cell(l)]) / e.c2;
        }
    }
}

// FS extraction (sediment_transport.cpp:167-201, :316-349): negative values
// become NDnen; the segment's end cells take the neighbouring value or keep
// FS as the reference's end tests decide.  xv() reads the solution at p.
template <int DIR, typename XV>
HD_DEV void fs_extract_at(const DevState &s, int line, int p, int f, int l, XV xv);
template <int DIR, typename XV>
HD_DEV void fs_extract_cell(const DevState &s, int line, int p, XV xv) {
    int k, f, l;
    if (locate<DIR>(s, line, p, k, f, l, false)) fs_extract_at<DIR>(s, line, p, f, l, xv);
}
// The same for a cell whose segment [f, l] is known.
template <int DIR, typename XV>
HD_DEV void fs_extract_at(const DevState &s, int line, int p, int f, int l, XV xv) {
    if (l - f < 2 || p < f + 1 || p > l - 1) return;
    bool ran1, ran2;
    wall_flags(s, DIR, line, f, l, ran1, ran2);
    const int W = s.P;
    auto cell = [&](int q) { return DIR == 1 ? q * W + line : line * W + q; };
    const double x = xv();
    const double out = x < 0 ? s.c.NDnen : x;
    const double *tvel = DIR == 1 ? s.t_u : s.t_v;
    s.tFS[cell(p)] = out;
    if (p == f + 1) {
        if ((ran1) || (tvel[cell(f)] == 0))
            s.tFS[cell(f)] = out;
        else
            s.tFS[cell(f)] = s.FS[cell(f)];
    }
    if (p == l - 1) {
        if ((ran2) || (tvel[cell(l)] == 0))
            s.tFS[cell(l)] = out;
        else
            s.tFS[cell(l)] = s.FS[cell(l)];
    }
}

// Cyclic reduction of a whole line in LDS, one workgroup per line; thread
// t owns rows q = t + k * blockDim.x, k < KPT.  The FP64 rows a thread needs again (its velocity row for the
// recovery, its reduced z row for the residual) stay in its registers, so
// each coefficient is read from HBM once and each unknown written once.
template <typename T, bool REFINE, bool INTERLEAVED, int KPT>
__global__ void __launch_bounds__(512) k_cr_reg(const DevState *__restrict__ S, SweepBufs Bin, int start, int end,
                                           int nrows) {
    const SweepBufs B = Bin.scen(blockIdx.y);
    extern __shared__ double smem_reg[];
    const int line = start + blockIdx.x;
    if (line > end) return;
    const int ld = B.ld;
    const double *__restrict__ A = B.A + (size_t)line * ld;
    const double *__restrict__ Bv = B.B + (size_t)line * ld;
    const double *__restrict__ C = B.C + (size_t)line * ld;
    const double *__restrict__ D = B.D + (size_t)line * ld;
    double *__restrict__ X = B.X + (size_t)line * ld;
    const int n = INTERLEAVED ? nrows / 2 : nrows;
    T *base = reinterpret_cast<T *>(smem_reg);
    CRShared<T> sh{base, base + n, base + 2 * n, base + 3 * n, base + 4 * n, base + 5 * n};
    const int tid = threadIdx.x, nt = blockDim.x;
    Row red[KPT], vel[KPT];
    double zf[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const int q = tid + k * nt;
        if (q >= n) continue;
        Row r;
        if (INTERLEAVED) {
            r = reduce_z_row(A, Bv, C, D, q, nrows);
            const int rv = 2 * q + 1;
            vel[k] = rv < nrows ? Row{A[rv], Bv[rv], C[rv], D[rv]} : Row{0.0, 1.0, 0.0, 0.0};
        } else {
            r = Row{A[q], Bv[q], C[q], D[q]};
        }
        sh.a[q] = (T)r.a;
        sh.b[q] = (T)r.b;
        sh.c[q] = (T)r.c;
        sh.d[q] = (T)r.d;
        red[k] = r;
    }
    __syncthreads();
    cr_solve<T>(sh, n, REFINE);
    if (REFINE) {
        // residual of the FP64 system at the T solution, solved again with
        // the stored multipliers
        double rr[KPT];
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int q = tid + k * nt;
            if (q >= n) continue;
            const double xq = (double)sh.d[q];
            zf[k] = xq;
            double v = red[k].d - red[k].b * xq;
            if (q > 0) v = v - red[k].a * (double)sh.d[q - 1];
            if (q + 1 < n) v = v - red[k].c * (double)sh.d[q + 1];
            rr[k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int q = tid + k * nt;
            if (q < n) sh.d[q] = (T)rr[k];
        }
        __syncthreads();
        cr_replay<T>(sh, n);
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int q = tid + k * nt;
            if (q < n) zf[k] = zf[k] + (double)sh.d[q];
        }
    } else {
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int q = tid + k * nt;
            if (q < n) zf[k] = (double)sh.d[q];
        }
    }
    if (!INTERLEAVED) {
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int q = tid + k * nt;
            if (q < n) X[q] = zf[k];
        }
        return;
    }
    // FP64 z of the line in LDS (over a/b, free after the solve; with T =
    // double sh.d already holds it) for the neighbours' velocity recovery
    const double *zs = reinterpret_cast<const double *>(sh.d);
    if (REFINE) {
        double *zw = smem_reg;
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int q = tid + k * nt;
            if (q < n) zw[q] = zf[k];
        }
        zs = zw;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const int q = tid + k * nt;
        if (q >= n) continue;
        const int r = 2 * q + 1;
        X[2 * q] = zf[k];
        if (r < nrows) {
            double v = vel[k].d;
            if (vel[k].a != 0.0) v = v - vel[k].a * zf[k];
            if (vel[k].c != 0.0 && r + 1 < nrows) v = v - vel[k].c * zs[q + 1];
            X[r] = v / vel[k].b;
        }
    }
}

// ---------------------------------------------------------------------------
// Partition solver, one workgroup per line ([line][rho]).
//
// Thread t owns the K consecutive unknowns q = tK .. tK+K-1 of the line's
// system (for hydraulics the z-only system, its rows reduced in FP64 as in
// reduce_z_vals).  It eliminates its first K-1 unknowns in registers
// (Thomas), leaving each as
//     x_k = y_k + v_k * L_{t-1} + w_k * L_t,
// where L_t is the thread's last unknown.  Substituted into the last row,
// this gives one row per thread in the L: a tridiagonal system (a Schur
// complement of the line's, so diagonally dominant like it) solved by
// parallel cyclic reduction in LDS, after which every x_k follows.  Rows pass
// through LDS on the way in and out so that global accesses stay coalesced
// while each thread works on a contiguous chunk.  With REFINE the T
// solution is corrected once by solving for the FP64 residual with the
// same elimination (den, cp, v, w and the reduced system's multipliers).
// Row lines (the VZ sweep, FS along i; S, the time and the clock flag are
// used only then): with ASSEMBLE the rows are formed from the state inside
// the kernel instead of read from B.A..D, and with EXTRACT the results
// (t_z and t_v, or tFS) are extracted from the LDS copy of the solution
// instead of X being written, so the whole sweep is this one kernel.
// Needs blockDim.x * K >= the number of unknowns.
// ---------------------------------------------------------------------------
template <typename T, bool REFINE, bool INTERLEAVED, int K, bool EXTRACT, bool ASSEMBLE = false>
__global__ void __launch_bounds__(512) k_part(SweepBufs Bin, int start, int end, int nrows,
                                              const DevState *__restrict__ S, double t_in, double s_start,
                                              int from_clock) {
    static_assert(K >= 2, "a chunk needs at least one interior unknown");
    static_assert(!ASSEMBLE || EXTRACT, "assembly is that of a row-line sweep, which extracts too");
    constexpr int R = INTERLEAVED ? 2 * K : K;  // rows per thread
    const SweepBufs B = Bin.scen(blockIdx.y);
    extern __shared__ double smem_part[];
    const int line = start + blockIdx.x;
    if (line > end) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = INTERLEAVED ? nrows / 2 : nrows;
    const size_t lo = (size_t)line * B.ld;
    // LDS.  The staging buffer keeps row rho at rho + rho / R, so that the R
    // rows of one thread are read with few bank conflicts.  The exchange
    // arrays and the reduced system overlap it: they are used only between
    // the inbound and the outbound staging, with barriers on both sides.
    double *stage = smem_part;
    T *ey = reinterpret_cast<T *>(smem_part);           // y_0 of each thread
    T *ev = reinterpret_cast<T *>(smem_part + nt);      // v_0
    T *ew = reinterpret_cast<T *>(smem_part + 2 * nt);  // w_0
    double *ez = smem_part + 3 * nt;                    // x_0 in FP64
    const PCRShared<T> sh{reinterpret_cast<T *>(smem_part + 4 * nt), nt};

    // Inbound: all four arrays are read coalesced (element tid + m * nt)
    // into registers first, then go through LDS one after the other, so
    // the later arrays' loads are in flight while the first ones are
    // staged.  The loads are unconditional (rows past nrows read row
    // nrows - 1 and are replaced by identity rows), which lets the
    // compiler wait for each array's loads alone.
    double own[4][R], prv[4];
    double ld4[4][R];
    // segments of the positions the thread assembles and extracts (row
    // lines: p = tid + m * nt + 2)
    int kk[K], ff[K], ll[K];
    if constexpr (ASSEMBLE && !INTERLEAVED) {
        // FS rows along i assembled here (as k_fs_assemble<0, 1>): position
        // p = tid + m * nt + 2 gives row p - 2, kept as ld4[.][m]
        const DevState &s = S[blockIdx.y];
        const double t = from_clock ? s.clock[0] : t_in;
        row_segments<K>(s, line, tid, nt, kk, ff, ll);
        bool r1[K], r2[K];  // the segments' wall tests, their loads in one round too
#pragma unroll
        for (int m = 0; m < K; m++) wall_flags(s, 0, line, ff[m], ll[m], r1[m], r2[m]);
#pragma unroll
        for (int m = 0; m < K; m++) {
            const int p = tid + m * nt + 2;
            Row r = identity_row();
            const int k = kk[m], f = ff[m], l = ll[m];
            if (p <= s.M) {
                if (k >= 0 && l - f >= 2) {
                    const bool ran1 = r1[m], ran2 = r2[m];
                    if (p >= f + 1 && p <= l - 1) {
                        fs_row_i_fast(s, line, p, f, l, ran1, ran2, t, s_start, r.a, r.b, r.c, r.d);
                        if (p == f + 1) r.a = 0.0;
                        if (p == l - 1) r.c = 0.0;
                    }
                    if (p == f) {
                        B.sys0[(size_t)line * s.capH + k] = f + 1 - 2;
                        B.sysn[(size_t)line * s.capH + k] = l - f - 2;
                    }
                } else if (k >= 0 && p == f) {
                    B.sys0[(size_t)line * s.capH + k] = -1;
                    B.sysn[(size_t)line * s.capH + k] = -1;
                }
            }
            ld4[0][m] = r.a; ld4[1][m] = r.b; ld4[2][m] = r.c; ld4[3][m] = r.d;
        }
    } else if constexpr (ASSEMBLE) {
        // VZ rows assembled here (as k_hyd_assemble<0, 1>): position
        // q = tid + m * nt gives rows 2q (z) and 2q + 1 (velocity), kept as
        // ld4[.][2m] and ld4[.][2m + 1]; positions past the line give
        // identity rows
        const DevState &s = S[blockIdx.y];
        const double f4 = 2 * s.c.g * s.c.dTchia2dY;
        row_segments<K>(s, line, tid, nt, kk, ff, ll);
#pragma unroll
        for (int m = 0; m < K; m++) {
            const int p = tid + m * nt + 2;
            Row rz = identity_row(), ru = identity_row();
            const int k = kk[m], f = ff[m], l = ll[m];
            if (p <= s.M && k >= 0) {
                SegCtx sc;
                hyd_rows<0, true>(s, line, p, f, l, f4, rz, ru, sc);
                if (p == f) {
                    B.sys0[(size_t)line * s.capH + k] = sc.rho0;
                    B.sysn[(size_t)line * s.capH + k] = sc.g.sn;
                }
            }
            ld4[0][2 * m] = rz.a; ld4[1][2 * m] = rz.b; ld4[2][2 * m] = rz.c; ld4[3][2 * m] = rz.d;
            ld4[0][2 * m + 1] = ru.a; ld4[1][2 * m + 1] = ru.b; ld4[2][2 * m + 1] = ru.c; ld4[3][2 * m + 1] = ru.d;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const double *src = a == 0 ? B.A : (a == 1 ? B.B : (a == 2 ? B.C : B.D));
            const double fill = a == 1 ? 1.0 : 0.0;
#pragma unroll
            for (int m = 0; m < R; m++) {
                const int e = tid + m * nt;
                const double v = src[lo + min(e, nrows - 1)];
                ld4[a][m] = e < nrows ? v : fill;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int m = 0; m < R; m++) {
            const int e = ASSEMBLE && INTERLEAVED ? 2 * (tid + (m / 2) * nt) + (m & 1) : tid + m * nt;
            stage[e + e / R] = ld4[a][m];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = tid * R + r;
            own[a][r] = stage[e + e / R];
        }
        if (INTERLEAVED) {
            const int e = tid * R - 1;
            prv[a] = tid > 0 ? stage[e + e / R] : (a == 1 ? 1.0 : 0.0);
        }
        __syncthreads();
    }

    // The thread's K rows (FP64) and, for hydraulics, its velocity rows.
    Row red[K], vel[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = tid * K + k;
        if (INTERLEAVED) {
            const Row m1 = k == 0 ? Row{prv[0], prv[1], prv[2], prv[3]}
                                  : Row{own[0][2 * k - 1], own[1][2 * k - 1], own[2][2 * k - 1], own[3][2 * k - 1]};
            const Row zr{own[0][2 * k], own[1][2 * k], own[2][2 * k], own[3][2 * k]};
            const Row p1{own[0][2 * k + 1], own[1][2 * k + 1], own[2][2 * k + 1], own[3][2 * k + 1]};
            // rows 2q + 1 < nrows exactly when q < n (nrows is even)
            red[k] = q < n ? reduce_z_vals(m1, zr, p1, q > 0, true) : identity_row();
            vel[k] = p1;
        } else {
            Row r{own[0][k], own[1][k], own[2][k], own[3][k]};
            if (q == 0) r.a = 0.0;
            if (q + 1 >= n) r.c = 0.0;
            red[k] = q < n ? r : identity_row();
        }
    }

    // Local elimination of the interior unknowns 0..K-2.
    T la[K], lc[K], den[K], cp[K], yv[K], vv[K], wv[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        la[k] = (T)red[k].a;
        lc[k] = (T)red[k].c;
    }
    den[0] = (T)red[0].b;
    cp[0] = lc[0] / den[0];
    yv[0] = (T)red[0].d / den[0];
    vv[0] = -la[0] / den[0];
#pragma unroll
    for (int k = 1; k < K - 1; k++) {
        den[k] = (T)red[k].b - la[k] * cp[k - 1];
        cp[k] = lc[k] / den[k];
        yv[k] = ((T)red[k].d - la[k] * yv[k - 1]) / den[k];
        vv[k] = -(la[k] * vv[k - 1]) / den[k];
    }
    wv[K - 2] = -cp[K - 2];
#pragma unroll
    for (int k = K - 3; k >= 0; k--) {
        yv[k] = yv[k] - cp[k] * yv[k + 1];
        vv[k] = vv[k] - cp[k] * vv[k + 1];
        wv[k] = -(cp[k] * wv[k + 1]);
    }

    // Row of L_t: a x_{K-2} + b L_t + c x'_0 = d, with x'_0 the next
    // thread's first unknown (zero past the last thread).
    ey[tid] = yv[0];
    ev[tid] = vv[0];
    ew[tid] = wv[0];
    __syncthreads();
    {
        const bool nb = tid + 1 < nt;
        const T yn = nb ? ey[tid + 1] : T(0), vn = nb ? ev[tid + 1] : T(0), wn = nb ? ew[tid + 1] : T(0);
        const T aK = la[K - 1], cK = lc[K - 1];
        pcr_solve<T>(sh, aK * vv[K - 2], (T)red[K - 1].b + aK * wv[K - 2] + cK * vn, cK * wn,
                     (T)red[K - 1].d - aK * yv[K - 2] - cK * yn, REFINE);
    }

    double zf[K];
    {
        const T Lt = sh.x()[tid], Lm = tid > 0 ? sh.x()[tid - 1] : T(0);
#pragma unroll
        for (int k = 0; k < K - 1; k++) zf[k] = (double)(yv[k] + vv[k] * Lm + wv[k] * Lt);
        zf[K - 1] = (double)Lt;
    }
    if (REFINE) {
        // FP64 residual of the thread's rows at the T solution ...
        ez[tid] = zf[0];
        __syncthreads();
        const double zn = tid + 1 < nt ? ez[tid + 1] : 0.0;
        const double zm = tid > 0 ? (double)sh.x()[tid - 1] : 0.0;
        T rv[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = tid * K + k;
            const double xl = k == 0 ? zm : zf[k - 1];
            const double xr = k == K - 1 ? zn : zf[k + 1];
            double v = red[k].d - red[k].b * zf[k];
            if (q > 0) v = v - red[k].a * xl;
            if (q + 1 < n) v = v - red[k].c * xr;
            rv[k] = (T)v;
        }
        // ... solved with the same elimination (v_k, w_k do not change).
        // ey is free again: every read of it happened before one of the
        // barriers above.
        T yr[K];
        yr[0] = rv[0] / den[0];
#pragma unroll
        for (int k = 1; k < K - 1; k++) yr[k] = (rv[k] - la[k] * yr[k - 1]) / den[k];
#pragma unroll
        for (int k = K - 3; k >= 0; k--) yr[k] = yr[k] - cp[k] * yr[k + 1];
        ey[tid] = yr[0];
        __syncthreads();
        const T yn = tid + 1 < nt ? ey[tid + 1] : T(0);
        pcr_replay<T>(sh, rv[K - 1] - la[K - 1] * yr[K - 2] - lc[K - 1] * yn);
        const T Ct = sh.x()[tid], Cm = tid > 0 ? sh.x()[tid - 1] : T(0);
#pragma unroll
        for (int k = 0; k < K - 1; k++) zf[k] = zf[k] + (double)(yr[k] + vv[k] * Cm + wv[k] * Ct);
        zf[K - 1] = zf[K - 1] + (double)Ct;
    }

    double out[R];
    if (INTERLEAVED) {
        // velocities from their own FP64 rows; the z after the chunk is the
        // next thread's first (ez was last read before a barrier above)
        ez[tid] = zf[0];
        __syncthreads();
        const double zn = tid + 1 < nt ? ez[tid + 1] : 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int r = 2 * (tid * K + k) + 1;
            out[2 * k] = zf[k];
            double v = vel[k].d;
            if (vel[k].a != 0.0) v = v - vel[k].a * zf[k];
            const double zr = k == K - 1 ? zn : zf[k + 1];
            if (vel[k].c != 0.0 && r + 1 < nrows) v = v - vel[k].c * zr;
            out[2 * k + 1] = v / vel[k].b;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = zf[k];
    }
    __syncthreads();  // the exchange and reduced-system arrays are dead from here
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int e = tid * R + r;
        stage[e + e / R] = out[r];
    }
    __syncthreads();
    if (EXTRACT) {
        // Extraction of a row line (the VZ sweep, or FS along i: writes
        // along j coalesce) for every cell of the row, from the solution
        // staged in LDS: the solution never goes to HBM, and no extraction
        // launch follows
        // (the positions are those of the assembly; their segments are
        // looked up again, which costs less than holding them through the
        // solve, where registers bound the workgroups per CU)
        const DevState &s = S[blockIdx.y];
        row_segments<K>(s, line, tid, nt, kk, ff, ll);
#pragma unroll
        for (int m = 0; m < K; m++) {
            const int p = 2 + tid + m * nt;
            if (p > s.M || kk[m] < 0) continue;
            if constexpr (INTERLEAVED)
                hyd_extract_at<0>(s, line, p, ff[m], ll[m], [&](int rho) { return stage[rho + rho / R]; });
            else
                fs_extract_at<0>(s, line, p, ff[m], ll[m], [&]() { return stage[(p - 2) + (p - 2) / R]; });
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < R; m++) {
        const int e = tid + m * nt;
        if (e < nrows) B.X[lo + e] = stage[e + e / R];
    }
}

template <int DIR, int LAYOUT>
__global__ void k_hyd_extract(const DevState *__restrict__ S, SweepBufs Bin, int start, int end) {
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int line = DIR == 1 ? j : i;
    const int p = DIR == 1 ? i : j;
    const int L = DIR == 1 ? s.N : s.M;
    if (line < start || line > end || p < 2 || p > L) return;
    const int ld = B.ld;
    hyd_extract_cell<DIR>(s, line, p, [&](int rho) {
        return LAYOUT == ROW_MAJOR ? B.X[rm_idx(rho, line, ld)] : B.X[lm_idx(rho, line, ld)];
    });
}

template <int DIR, int LAYOUT>
__global__ void k_fs_extract(const DevState *__restrict__ S, SweepBufs Bin, int start, int end) {
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int line = DIR == 1 ? j : i;
    const int p = DIR == 1 ? i : j;
    const int L = DIR == 1 ? s.N : s.M;
    if (line < start || line > end || p < 2 || p > L) return;
    const int ld = B.ld;
    fs_extract_cell<DIR>(s, line, p, [&]() {
        return LAYOUT == ROW_MAJOR ? B.X[rm_idx(p - 2, line, ld)] : B.X[lm_idx(p - 2, line, ld)];
    });
}

// Column lines from [line][rho]: a tile of 64 lines x 8 cells along i
// (i0 = 2 + 8 * blockIdx.y) first reads the rows its cells need, coalesced
// along each line, into LDS: rows 2(i0-3) .. 2(i0+5)+1 for hydraulics (the
// cells' own two rows and those of the cell before, which the segment-end
// formulas read), rows i0-2 .. i0+5 for FS.  Rows outside that window are
// read from global memory.  Same results as k_hyd_extract / k_fs_extract
// with DIR = 1.
constexpr int XT_ROWS = 18;

__global__ void __launch_bounds__(512) k_hyd_extract_cols(const DevState *__restrict__ S, SweepBufs Bin, int start,
                                                          int end, int nrows) {
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    __shared__ double xs[XT_ROWS][64 + 1];
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * 64 + tx, ld = B.ld;
    const int j0 = blockIdx.x * 64, i0 = 2 + blockIdx.y * 8;
    const int rb = 2 * (i0 - 3);  // first row of the window (may be negative)
    for (int e = t; e < 64 * XT_ROWS; e += 512) {
        const int ln = e / XT_ROWS, r = e % XT_ROWS;  // r runs fastest: contiguous along the line
        const int jj = j0 + ln, rho = rb + r;
        if (jj >= start && jj <= end && rho >= 0 && rho < nrows) xs[r][ln] = B.X[lm_idx(rho, jj, ld)];
    }
    __syncthreads();
    const int j = j0 + tx, i = i0 + ty;
    if (j < start || j > end || i > s.N) return;
    hyd_extract_cell<1>(s, j, i, [&](int rho) {
        const int r = rho - rb;
        return (r >= 0 && r < XT_ROWS) ? xs[r][tx] : B.X[lm_idx(rho, j, ld)];
    });
}

__global__ void __launch_bounds__(512) k_fs_extract_cols(const DevState *__restrict__ S, SweepBufs Bin, int start,
                                                         int end) {
    const SweepBufs B = Bin.scen(blockIdx.z);
    const DevState &s = S[blockIdx.z];
    __shared__ double xs[8][64 + 1];
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * 64 + tx, ld = B.ld;
    const int j0 = blockIdx.x * 64, i0 = 2 + blockIdx.y * 8;
    {
        const int ln = t / 8, r = t % 8;  // 512 threads: one row each
        const int jj = j0 + ln, ci = i0 + r;
        if (jj >= start && jj <= end && ci <= s.N) xs[r][ln] = B.X[lm_idx(ci - 2, jj, ld)];
    }
    __syncthreads();
    const int j = j0 + tx, i = i0 + ty;
    if (j < start || j > end || i > s.N) return;
    fs_extract_cell<1>(s, j, i, [&]() { return xs[ty][tx]; });
}

// ---------------------------------------------------------------------------
// Host sequencing of one sweep
// ---------------------------------------------------------------------------
static dim3 tile_grid(const DevState &s, int n) { return dim3((s.M + 3 + 63) / 64, (s.N + 3 + 7) / 8, n); }

int sweep_rows(const DevState &s, int dir, bool hyd) {
    const int L = dir == 1 ? s.N : s.M;
    return hyd ? 2 * (L - 1) : (L - 1);
}

// Assembly of one sweep's rows (column lines in [line][rho] take the
// LDS-transposing tile kernels).
template <int DIR, int LAYOUT>
static int launch_assembly(hd_ctx *ctx, const DevState *S, const DevState &s, const SweepBufs &B, int n, bool hyd,
                           int start, int end, double t, double s_start, int from_clock) {
    hipStream_t st = ctx->stream;
    if (DIR == 1 && LAYOUT == LINE_MAJOR) {
        constexpr int TRH = 8 * COLS_HYD_CPT, TRF = 8 * COLS_FS_CPT;
        const int cells = s.N - 1, gx = (s.M + 3 + 63) / 64;
        if (hyd)
            LAUNCH(ctx, "k_hyd_assemble<uz>",
                   (k_hyd_assemble_cols<COLS_HYD_CPT><<<dim3(gx, (cells + TRH - 1) / TRH, n), dim3(64, 8), 0, st>>>(
                       S, B, start, end)));
        else
            LAUNCH(ctx, "k_fs_assemble<j>",
                   (k_fs_assemble_cols<COLS_FS_CPT><<<dim3(gx, (cells + TRF - 1) / TRF, n), dim3(64, 8), 0, st>>>(
                       S, B, start, end, t, s_start, from_clock)));
        return HD_OK;
    }
    const dim3 tg = tile_grid(s, n);
    if (hyd)
        LAUNCH(ctx, DIR ? "k_hyd_assemble<uz>" : "k_hyd_assemble<vz>",
               (k_hyd_assemble<DIR, LAYOUT><<<tg, dim3(64, 8), 0, st>>>(S, B, start, end)));
    else
        LAUNCH(ctx, DIR ? "k_fs_assemble<j>" : "k_fs_assemble<i>",
               (k_fs_assemble<DIR, LAYOUT><<<tg, dim3(64, 8), 0, st>>>(S, B, start, end, t, s_start, from_clock)));
    return HD_OK;
}

// Lets a kernel use more than the default 64 KiB of dynamic LDS (gfx950
// has 160 KiB per CU).
template <typename K>
static int allow_lds(hd_ctx *ctx, K kernel, size_t bytes) {
    if (bytes <= 65536) return HD_OK;
    hipError_t e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) {
        ctx->err = std::string("hipFuncSetAttribute(LDS): ") + hipGetErrorString(e);
        return HD_E_HIP;
    }
    return HD_OK;
}

// LDS of k_part: the staging buffer, or the exchange arrays plus the
// reduced system, whichever is larger.
template <typename T, bool REFINE>
static size_t part_lds_bytes(int nt, int rows_per_thread) {
    const size_t stage = (size_t)nt * (rows_per_thread + 1) * sizeof(double);
    const size_t other = (size_t)nt * 4 * sizeof(double) +
                         (size_t)nt * (10 + (REFINE ? 2 * pcr_steps(nt) : 0)) * sizeof(T);
    return stage > other ? stage : other;
}

// Launch shape of a line solver for lines of nz unknowns: workgroup size
// and unknowns per thread (the template K).  Returns false, with the reason
// in ctx->err, when the line is too long for that solver.
static bool solver_shape(hd_ctx *ctx, int nz, int kind, int param, int &threads, int &kpt) {
    if (kind == HD_SOLVER_CR) {
        threads = param ? param : (nz > 1024 ? 512 : 256);
        const int need = (nz + threads - 1) / threads;
        if (need > 8) {
            ctx->err = "line of " + std::to_string(nz) + " unknowns exceeds cyclic reduction's 8 x " +
                       std::to_string(threads);
            return false;
        }
        kpt = need <= 2 ? 2 : (need <= 4 ? 4 : 8);
        return true;
    }
    kpt = param ? param : (nz <= 4 * 512 ? 4 : 8);
    const int chunks = (nz + kpt - 1) / kpt;
    threads = std::max(64, (chunks + 63) / 64 * 64);
    if (threads > 512) {
        ctx->err = "line of " + std::to_string(nz) + " unknowns exceeds the partition solver's 512 x " +
                   std::to_string(kpt);
        return false;
    }
    return true;
}

// Launches one fast-form line solver over `lines` lines of nscen scenarios.
// extract (partition solver, row lines): the solver also extracts t_z and
// t_v (hydraulics) or tFS (FS) from its LDS copy of the solution instead of
// writing X.
template <typename T, bool REFINE, bool IL>
static int launch_solver(hd_ctx *ctx, const DevState *S, const SweepBufs &B, int nscen, int kind, int threads,
                         int kpt, int start, int end, int nrows, const char *name, bool extract, bool assemble,
                         double t, double s_start) {
    const int from_clock = ctx->capturing ? 1 : 0;
    hipStream_t st = ctx->stream;
    const int lines = end - start + 1;
    int rc = HD_OK;
    if (kind == HD_SOLVER_CR) {
        const int nz = IL ? nrows / 2 : nrows;
        const size_t sm = (REFINE ? 6 : 4) * (size_t)nz * sizeof(T);
        auto go = [&](auto kern) {
            rc = allow_lds(ctx, kern, sm);
            if (rc) return;
            LAUNCH_RC(ctx, rc, name, (kern<<<dim3(lines, nscen), threads, sm, st>>>(S, B, start, end, nrows)));
        };
        if (kpt == 2) go(k_cr_reg<T, REFINE, IL, 2>);
        else if (kpt == 4) go(k_cr_reg<T, REFINE, IL, 4>);
        else go(k_cr_reg<T, REFINE, IL, 8>);
    } else {
        const size_t sm = part_lds_bytes<T, REFINE>(threads, IL ? 2 * kpt : kpt);
        auto go = [&](auto kern) {
            rc = allow_lds(ctx, kern, sm);
            if (rc) return;
            LAUNCH_RC(ctx, rc, name,
                      (kern<<<dim3(lines, nscen), threads, sm, st>>>(B, start, end, nrows, S, t, s_start,
                                                                      from_clock)));
        };
        if (extract && assemble) {
            if (kpt == 2) go(k_part<T, REFINE, IL, 2, true, true>);
            else if (kpt == 4) go(k_part<T, REFINE, IL, 4, true, true>);
            else go(k_part<T, REFINE, IL, 8, true, true>);
        } else if (extract) {
            if (kpt == 2) go(k_part<T, REFINE, IL, 2, true>);
            else if (kpt == 4) go(k_part<T, REFINE, IL, 4, true>);
            else go(k_part<T, REFINE, IL, 8, true>);
        } else {
            if (kpt == 2) go(k_part<T, REFINE, IL, 2, false>);
            else if (kpt == 4) go(k_part<T, REFINE, IL, 4, false>);
            else go(k_part<T, REFINE, IL, 8, false>);
        }
    }
    return rc;
}

// Stage names of the fast-form solvers, [kind == CR][f32][dir][hydraulic].
static const char *solver_name(int kind, bool f32, int dir, bool hyd) {
    static const char *names[2][2][2][2] = {
        {{{"k_part<f64,fs_i>", "k_part<f64,vz>"}, {"k_part<f64,fs_j>", "k_part<f64,uz>"}},
         {{"k_part<f32r,fs_i>", "k_part<f32r,vz>"}, {"k_part<f32r,fs_j>", "k_part<f32r,uz>"}}},
        {{{"k_cr_reg<f64,fs_i>", "k_cr_reg<f64,vz>"}, {"k_cr_reg<f64,fs_j>", "k_cr_reg<f64,uz>"}},
         {{"k_cr_reg<f32r,fs_i>", "k_cr_reg<f32r,vz>"}, {"k_cr_reg<f32r,fs_j>", "k_cr_reg<f32r,uz>"}}}};
    return names[kind == HD_SOLVER_CR][f32][dir][hyd];
}

template <int LAYOUT>
static int seg_solve(hd_ctx *ctx, const DevState *S, const SweepBufs &B, int n, int dir, bool hyd, int start,
                     int end);

static int run_solver(hd_ctx *ctx, const DevState *S, const SweepBufs &B, int nscen, int dir, bool hyd, int kind,
                      int param, int start, int end, int nrows, bool extract = false, bool assemble = false,
                      double t = 0.0, double s_start = 0.0) {
    // FP64 Thomas with one lane per system (the fast FP64 form's extra
    // candidate; it reads the systems' rows as assembled, in FP64)
    if (kind == HD_SOLVER_SEGMENTS) return seg_solve<LINE_MAJOR>(ctx, S, B, nscen, dir, hyd, start, end);
    const int nz = hyd ? nrows / 2 : nrows;
    if (kind == HD_SOLVER_DEFAULT) kind = HD_SOLVER_PARTITION;
    int threads, kpt;
    if (!solver_shape(ctx, nz, kind, param, threads, kpt)) return HD_E_UNSUPPORTED;
    const bool f32 = ctx->cfg.form == HD_FORM_FP32_REFINE;
    const char *name = solver_name(kind, f32, dir, hyd);
    extract = extract && dir == 0 && kind == HD_SOLVER_PARTITION;
    assemble = assemble && extract;
    if (f32)
        return hyd ? launch_solver<float, true, true>(ctx, S, B, nscen, kind, threads, kpt, start, end, nrows, name,
                                                      extract, assemble, t, s_start)
                   : launch_solver<float, true, false>(ctx, S, B, nscen, kind, threads, kpt, start, end, nrows, name,
                                                       extract, assemble, t, s_start);
    return hyd ? launch_solver<double, false, true>(ctx, S, B, nscen, kind, threads, kpt, start, end, nrows, name,
                                                    extract, assemble, t, s_start)
               : launch_solver<double, false, false>(ctx, S, B, nscen, kind, threads, kpt, start, end, nrows, name,
                                                     extract, assemble, t, s_start);
}

// Whether the solver of a row-line sweep (VZ, or FS along i) extracts the
// solution itself (the partition solver does; cyclic reduction leaves it
// to k_hyd_extract / k_fs_extract).
static bool solver_extracts(const hd_ctx *ctx, bool hyd) {
    const int kind = ctx->solver_kind[0][hyd ? 1 : 0];
    return kind == HD_SOLVER_DEFAULT || kind == HD_SOLVER_PARTITION;
}

// ---------------------------------------------------------------------------
// Thomas with one lane per system (HD_SOLVER_SEGMENTS, the FP64 forms).  The
// systems of a line are independent (identity rows between them), and
// k_thomas's walk along a line carries values only across rows outside
// every system, which no system reads: its first row divides by its own
// diagonal, its last row ignores the value after it.  Solving each system
// on a lane of its own therefore gives the same bits, with the serial
// chain cut from a line's length to its longest system's and every system
// of the sweep in flight at once.  k_seg_list lists the sweep's systems
// (index line * cap + k into sys0 / sysn) in line order, one workgroup
// per scenario (a workgroup-wide prefix of per-thread counts); k_thomas_seg
// then gives list entry e to lane e (and e + stride, ... when the grid is
// smaller than the list).  A system's row place (first, interior, last,
// only) follows from its row index, so the row-code bytes are not read.
// LAYOUT: ROW_MAJOR (the plain form's [rho][line]) or LINE_MAJOR (the
// fast FP64 form's [line][rho], where the same FP64 Thomas is one more
// candidate of the tuner).
// ---------------------------------------------------------------------------
constexpr int SEGLIST_THREADS = 1024;

__global__ void __launch_bounds__(SEGLIST_THREADS) k_seg_list(const DevState *__restrict__ S, SweepBufs Bin, int dir,
                                                             int start, int end) {
    const SweepBufs B = Bin.scen(blockIdx.x);
    const DevState &s = S[blockIdx.x];
    const int cap = dir == 1 ? s.capV : s.capH;
    const int *__restrict__ cnt = dir == 1 ? s.segV_n : s.segH_n;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int per = (end - start + 1 + nt - 1) / nt;  // consecutive lines per thread
    const int l0 = start + tid * per, l1 = min(end + 1, l0 + per);
    auto live = [&](size_t o) { return B.sys0[o] >= 0 && B.sysn[o] >= 0; };
    int mine = 0;
    for (int L = l0; L < l1; L++)
        for (int k = 0; k < cnt[L]; k++) mine += live((size_t)L * cap + k) ? 1 : 0;
    __shared__ int sc[SEGLIST_THREADS];
    sc[tid] = mine;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {  // inclusive scan
        const int v = tid >= off ? sc[tid - off] : 0;
        __syncthreads();
        sc[tid] += v;
        __syncthreads();
    }
    int pos = sc[tid] - mine;
    for (int L = l0; L < l1; L++)
        for (int k = 0; k < cnt[L]; k++) {
            const size_t o = (size_t)L * cap + k;
            if (live(o)) B.segl[pos++] = (int)o;
        }
    if (tid == nt - 1) B.segcnt[0] = sc[tid];
}

template <int LAYOUT>
__global__ void __launch_bounds__(64) k_thomas_seg(SweepBufs Bin, const DevState *__restrict__ S, int dir) {
    const SweepBufs B = Bin.scen(blockIdx.y);
    const int cap = dir == 1 ? S[blockIdx.y].capV : S[blockIdx.y].capH;
    const int total = B.segcnt[0], ld = B.ld, stride = gridDim.x * 64;
    // element (rho, line) of the layout: base + line * lstep + rho * rstep
    const size_t lstep = LAYOUT == ROW_MAJOR ? 1 : (size_t)ld, rstep = LAYOUT == ROW_MAJOR ? (size_t)ld : 1;
    for (int e = blockIdx.x * 64 + threadIdx.x; e < total; e += stride) {
        const int o = B.segl[e];
        const int line = o / cap, r0 = B.sys0[o], sn = B.sysn[o], nr = sn + 1;
        const size_t lo = (size_t)line * lstep;
        const double *__restrict__ A = B.A + lo;
        const double *__restrict__ Bv = B.B + lo;
        const double *__restrict__ C = B.C + lo;
        const double *__restrict__ D = B.D + lo;
        double *__restrict__ Ap = B.Ap + lo;
        double *__restrict__ Bp = B.Bp + lo;
        double *__restrict__ X = B.X + lo;
        auto at = [&](int q) { return (size_t)(r0 + min(max(q, 0), sn)) * rstep; };
        auto code_of = [&](int q) { return sn == 0 ? 4 : (q == 0 ? 1 : (q == sn ? 3 : 2)); };
        auto load_row = [&](int q) {
            const size_t w = at(q);
            return ThomasRow{A[w], Bv[w], C[w], D[w], code_of(min(max(q, 0), sn))};
        };
        // forward, as k_thomas (TB rows of loads in flight ahead)
        double ApP = 0, BpP = 0;
        auto fwd = [&](const ThomasRow &w, int q) {
            const bool first = w.code == 1 || w.code == 4;
            const double den = first ? w.b : w.a * ApP + w.b;
            const double num = first ? w.d : w.d - (w.a * BpP);
            const double ap = (w.code == 1 || w.code == 2) ? -w.c / den : 0.0;
            const double bp = num / den;
            ApP = ap;
            BpP = bp;
            Ap[at(q)] = ap;
            Bp[at(q)] = bp;
        };
        const int nfull = nr / TB;
        ThomasRow cur[TB], nxt[TB];
#pragma unroll
        for (int q = 0; q < TB; q++) nxt[q] = load_row(q);
        for (int blk = 0; blk < nfull; blk++) {
#pragma unroll
            for (int q = 0; q < TB; q++) cur[q] = nxt[q];
#pragma unroll
            for (int q = 0; q < TB; q++) nxt[q] = load_row((blk + 1) * TB + q);
#pragma unroll
            for (int q = 0; q < TB; q++) fwd(cur[q], blk * TB + q);
        }
#pragma unroll
        for (int q = 0; q < TB; q++) {
            if (nfull * TB + q >= nr) break;
            fwd(nxt[q], nfull * TB + q);
        }
        // backward, as k_thomas
        double xnext = 0;
        auto bwd = [&](int q, double pa, double pb) {
            const int code = code_of(q);
            const double xv = (code == 3 || code == 4) ? pb : pb + (pa * xnext);
            xnext = xv;
            X[at(q)] = xv;
        };
        for (int q = sn; q >= nfull * TB; q--) bwd(q, Ap[at(q)], Bp[at(q)]);
        double pa[TB], pb[TB], qa[TB], qb[TB];
#pragma unroll
        for (int u = 0; u < TB; u++) {
            qa[u] = Ap[at(nfull * TB - 1 - u)];
            qb[u] = Bp[at(nfull * TB - 1 - u)];
        }
        for (int blk = nfull - 1; blk >= 0; blk--) {
#pragma unroll
            for (int u = 0; u < TB; u++) {
                pa[u] = qa[u];
                pb[u] = qb[u];
            }
#pragma unroll
            for (int u = 0; u < TB; u++) {
                qa[u] = Ap[at(blk * TB - 1 - u)];
                qb[u] = Bp[at(blk * TB - 1 - u)];
            }
#pragma unroll
            for (int u = 0; u < TB; u++) bwd(blk * TB + TB - 1 - u, pa[u], pb[u]);
        }
    }
}

// Thomas with one lane per system over a sweep's assembled systems in
// LAYOUT (8 lanes per line in the grid; lanes loop when a sweep has more
// systems).  B.ld is the layout's leading dimension.
template <int LAYOUT>
static int seg_solve(hd_ctx *ctx, const DevState *S, const SweepBufs &B, int n, int dir, bool hyd, int start,
                     int end) {
    hipStream_t st = ctx->stream;
    const int lines = end - start + 1;
    LAUNCH(ctx, "k_seg_list", k_seg_list<<<n, SEGLIST_THREADS, 0, st>>>(S, B, dir, start, end));
    const int waves = std::max(16, (lines * 8 + 63) / 64);
    LAUNCH(ctx,
           hyd ? (dir ? "k_thomas_seg<uz>" : "k_thomas_seg<vz>") : (dir ? "k_thomas_seg<fs_j>" : "k_thomas_seg<fs_i>"),
           k_thomas_seg<LAYOUT><<<dim3(waves, n), 64, 0, st>>>(B, S, dir));
    return HD_OK;
}

// The plain form's solve of one sweep: Thomas per line (k_thomas), or per
// system (kind HD_SOLVER_SEGMENTS).  B.ld is the ROW_MAJOR leading
// dimension.
template <int DIR>
static int plain_solve(hd_ctx *ctx, const DevState *S, const DevState &s, const SweepBufs &B, int n, bool hyd,
                       int start, int end, int nrows, int kind) {
    hipStream_t st = ctx->stream;
    const int lines = end - start + 1;
    if (kind == HD_SOLVER_SEGMENTS) return seg_solve<ROW_MAJOR>(ctx, S, B, n, DIR, hyd, start, end);
    LAUNCH(ctx, hyd ? (DIR ? "k_thomas<uz>" : "k_thomas<vz>") : (DIR ? "k_thomas<fs_j>" : "k_thomas<fs_i>"),
           k_thomas<<<dim3((lines + 63) / 64, n), 64, 0, st>>>(B, start, end, nrows));
    return HD_OK;
}

template <int DIR, int LAYOUT>
static int sweep_impl(hd_ctx *ctx, const DevState *S, const DevState &s, int n, bool hyd, int start, int end,
                      double t, double s_start) {
    hipStream_t st = ctx->stream;
    SweepBufs B = ctx->sweep;
    const int nrows = sweep_rows(s, DIR, hyd);
    B.ld = LAYOUT == ROW_MAJOR ? B.ld_rm[DIR] : B.ld_lm[DIR];
    // row-line sweeps (VZ, FS along i) are assembled inside the partition
    // solver
    const bool fused_asm = LAYOUT == LINE_MAJOR && DIR == 0 && solver_extracts(ctx, hyd);
    int rc = HD_OK;
    if (!fused_asm) rc = launch_assembly<DIR, LAYOUT>(ctx, S, s, B, n, hyd, start, end, t, s_start, ctx->capturing ? 1 : 0);
    if (rc) return rc;
    const int lines = end - start + 1;
    if (LAYOUT == ROW_MAJOR) {
        rc = plain_solve<DIR>(ctx, S, s, B, n, hyd, start, end, nrows, ctx->solver_kind[DIR][hyd ? 1 : 0]);
        if (rc) return rc;
    } else {
        const int h = hyd ? 1 : 0;
        const bool fused = DIR == 0 && solver_extracts(ctx, hyd);
        rc = run_solver(ctx, S, B, n, DIR, hyd, ctx->solver_kind[DIR][h], ctx->solver_param[DIR][h], start, end,
                        nrows, fused, fused_asm, t, s_start);
        if (rc) return rc;
        if (fused) return HD_OK;  // t_z, t_v (or tFS) written by the solver
    }
    if (DIR == 1 && LAYOUT == LINE_MAJOR) {
        const dim3 xg((s.M + 3 + 63) / 64, (s.N - 1 + 7) / 8, n);
        if (hyd)
            LAUNCH(ctx, "k_hyd_extract<uz>",
                   (k_hyd_extract_cols<<<xg, dim3(64, 8), 0, st>>>(S, B, start, end, nrows)));
        else
            LAUNCH(ctx, "k_fs_extract<j>", (k_fs_extract_cols<<<xg, dim3(64, 8), 0, st>>>(S, B, start, end)));
        return HD_OK;
    }
    const Launch2D g = grid2d(s, n);
    if (hyd)
        LAUNCH(ctx, DIR ? "k_hyd_extract<uz>" : "k_hyd_extract<vz>",
               (k_hyd_extract<DIR, LAYOUT><<<g.grid, g.block, 0, st>>>(S, B, start, end)));
    else
        LAUNCH(ctx, DIR ? "k_fs_extract<j>" : "k_fs_extract<i>",
               (k_fs_extract<DIR, LAYOUT><<<g.grid, g.block, 0, st>>>(S, B, start, end)));
    return HD_OK;
}

// Copies the fields a row-line sweep writes (t_z and t_v, or tFS) of every
// scenario into its keep0 / keep1, or back.
static int swap_sweep_outputs(hd_ctx *ctx, int n, bool hyd, bool save) {
    for (int k = 0; k < n; k++) {
        const DevState &sk = k == 0 ? ctx->hs : ctx->hss[k - 1];
        const size_t bytes = (size_t)sk.P * sk.H * sizeof(double);
        double *f0 = hyd ? sk.t_z : sk.tFS, *f1 = sk.t_v;
        double *a = sk.keep0, *b = sk.keep1;
        if (hipMemcpyAsync(save ? a : f0, save ? f0 : a, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            return HD_E_HIP;
        if (hyd && hipMemcpyAsync(save ? b : f1, save ? f1 : b, bytes, hipMemcpyDeviceToDevice, ctx->stream) !=
                       hipSuccess)
            return HD_E_HIP;
    }
    return HD_OK;
}

// Times every line-solver candidate on one sweep and keeps the fastest.
// Column-line sweeps: the system is assembled once from the current state
// and only the solve is timed (assembly and extraction are the same for
// every candidate).  Row-line sweeps: the partition solver assembles and
// extracts inside its kernel, so each candidate is timed as the whole
// sweep, with the fields the sweep writes saved before and put back after.
// Either way only sweep scratch keeps a trace: the state is untouched.
template <int DIR>
static int tune_impl(hd_ctx *ctx, const DevState *S, const DevState &s, int n, bool hyd, int start, int end,
                     double t, double s_start, std::string &report) {
    hipStream_t st = ctx->stream;
    SweepBufs B = ctx->sweep;
    const int nrows = sweep_rows(s, DIR, hyd);
    B.ld = B.ld_lm[DIR];
    const bool whole = DIR == 0;
    int arc = whole ? swap_sweep_outputs(ctx, n, hyd, true)
                    : launch_assembly<DIR, LINE_MAJOR>(ctx, S, s, B, n, hyd, start, end, t, s_start, 0);
    if (arc) return arc;
    const int h = hyd ? 1 : 0;
    const int kind0 = ctx->solver_kind[DIR][h], param0 = ctx->solver_param[DIR][h];
    auto candidate = [&](int kind, int param) {
        if (!whole) return run_solver(ctx, S, B, n, DIR, hyd, kind, param, start, end, nrows);
        ctx->solver_kind[DIR][h] = kind;
        ctx->solver_param[DIR][h] = param;
        return sweep_impl<DIR, LINE_MAJOR>(ctx, S, s, n, hyd, start, end, t, s_start);
    };
    const int nz = hyd ? nrows / 2 : nrows;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return HD_E_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return HD_E_HIP;
    }
    const bool prof = ctx->prof;
    ctx->prof = false;  // candidates are timed here, not as stages
    int best_kind = 0, best_param = 0, rc = HD_OK;
    float best_ms = 0;
    char buf[160];
    // (Thomas per system, FP64, only for the fast FP64 form)
    const int cands[7][2] = {{HD_SOLVER_PARTITION, 2}, {HD_SOLVER_PARTITION, 4}, {HD_SOLVER_PARTITION, 8},
                             {HD_SOLVER_CR, 128},      {HD_SOLVER_CR, 256},      {HD_SOLVER_CR, 512},
                             {HD_SOLVER_SEGMENTS, 0}};
    for (const auto &c : cands) {
        int threads, kpt;
        if (c[0] == HD_SOLVER_SEGMENTS) {
            if (ctx->cfg.form != HD_FORM_FP64_FAST) continue;
        } else if (!solver_shape(ctx, nz, c[0], c[1], threads, kpt)) {
            continue;  // too long for this candidate
        }
        rc = candidate(c[0], c[1]);  // warm-up
        if (!rc) rc = hipEventRecord(e0, st) == hipSuccess ? HD_OK : HD_E_HIP;
        for (int r = 0; r < 3 && !rc; r++) rc = candidate(c[0], c[1]);
        if (!rc) rc = hipEventRecord(e1, st) == hipSuccess ? HD_OK : HD_E_HIP;
        if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? HD_OK : HD_E_HIP;
        if (rc) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= 3;
        std::snprintf(buf, sizeof(buf), "%s\"%s%s/%s%d\": %.4f", report.size() > 1 ? ", " : "", DIR ? "uz" : "vz",
                      hyd ? "" : "_fs", c[0] == HD_SOLVER_CR ? "cr" : (c[0] == HD_SOLVER_SEGMENTS ? "seg" : "part"),
                      c[1], ms);
        report += buf;
        if (!best_kind || ms < best_ms) {
            best_kind = c[0];
            best_param = c[1];
            best_ms = ms;
        }
    }
    ctx->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->solver_kind[DIR][h] = kind0;
    ctx->solver_param[DIR][h] = param0;
    if (whole) {
        const int src = swap_sweep_outputs(ctx, n, hyd, false);
        if (!rc) rc = src;
        if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = HD_E_HIP;
    }
    if (rc) return rc;
    if (!best_kind) return HD_E_UNSUPPORTED;  // ctx->err names the line length
    ctx->solver_kind[DIR][hyd ? 1 : 0] = best_kind;
    ctx->solver_param[DIR][hyd ? 1 : 0] = best_param;
    return HD_OK;
}

// The plain form's tuning: Thomas per line against Thomas per system on
// the rows assembled from the current state.  Both write only sweep
// scratch and give the same bits, so the choice changes no result.
template <int DIR>
static int tune_plain(hd_ctx *ctx, const DevState *S, const DevState &s, int n, bool hyd, int start, int end,
                      double t, double s_start, std::string &report) {
    hipStream_t st = ctx->stream;
    SweepBufs B = ctx->sweep;
    const int nrows = sweep_rows(s, DIR, hyd);
    B.ld = B.ld_rm[DIR];
    int rc = launch_assembly<DIR, ROW_MAJOR>(ctx, S, s, B, n, hyd, start, end, t, s_start, 0);
    if (rc) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return HD_E_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return HD_E_HIP;
    }
    const bool prof = ctx->prof;
    ctx->prof = false;  // candidates are timed here, not as stages
    int best = HD_SOLVER_DEFAULT;
    float best_ms = 0;
    char buf[160];
    for (int kind : {HD_SOLVER_DEFAULT, HD_SOLVER_SEGMENTS}) {
        rc = plain_solve<DIR>(ctx, S, s, B, n, hyd, start, end, nrows, kind);  // warm-up
        if (!rc) rc = hipEventRecord(e0, st) == hipSuccess ? HD_OK : HD_E_HIP;
        for (int r = 0; r < 3 && !rc; r++) rc = plain_solve<DIR>(ctx, S, s, B, n, hyd, start, end, nrows, kind);
        if (!rc) rc = hipEventRecord(e1, st) == hipSuccess ? HD_OK : HD_E_HIP;
        if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? HD_OK : HD_E_HIP;
        if (rc) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= 3;
        std::snprintf(buf, sizeof(buf), "%s\"%s%s/thomas_%s\": %.4f", report.size() > 1 ? ", " : "", DIR ? "uz" : "vz",
                      hyd ? "" : "_fs", kind == HD_SOLVER_SEGMENTS ? "seg" : "line", ms);
        report += buf;
        if (kind == HD_SOLVER_DEFAULT || ms < best_ms) {
            best = kind;
            best_ms = ms;
        }
    }
    ctx->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    ctx->solver_kind[DIR][hyd ? 1 : 0] = best;
    ctx->solver_param[DIR][hyd ? 1 : 0] = 0;
    return HD_OK;
}

int tune_line_solvers(hd_ctx *ctx, const DevState *S, const DevState &s, int n, std::string &report) {
    const hd_options &o = ctx->ops;
    if (ctx->cfg.form == HD_FORM_FP64_PLAIN) {
        int rc = tune_plain<1>(ctx, S, s, n, true, 2, s.M, 0.0, 0.0, report);
        if (!rc) rc = tune_plain<0>(ctx, S, s, n, true, 2, s.N, 0.0, 0.0, report);
        if (!rc && ctx->cfg.sediment)
            rc = tune_plain<1>(ctx, S, s, n, false, 3, s.M - 1, o.sediment_start, o.bed_change_start, report);
        if (!rc && ctx->cfg.sediment)
            rc = tune_plain<0>(ctx, S, s, n, false, 3, s.N - 1, o.sediment_start, o.bed_change_start, report);
        return rc;
    }
    int rc = tune_impl<1>(ctx, S, s, n, true, 2, s.M, 0.0, 0.0, report);
    if (!rc) rc = tune_impl<0>(ctx, S, s, n, true, 2, s.N, 0.0, 0.0, report);
    if (!rc && ctx->cfg.sediment) rc = tune_impl<1>(ctx, S, s, n, false, 3, s.M - 1, o.sediment_start, o.bed_change_start, report);
    if (!rc && ctx->cfg.sediment) rc = tune_impl<0>(ctx, S, s, n, false, 3, s.N - 1, o.sediment_start, o.bed_change_start, report);
    return rc;
}

int solve_batch(hd_ctx *ctx, const double *dl, const double *d, const double *du, const double *b, double *x, int m,
                int nsys, int ld) {
    SweepBufs B{};
    B.A = const_cast<double *>(dl);
    B.B = const_cast<double *>(d);
    B.C = const_cast<double *>(du);
    B.D = const_cast<double *>(b);
    B.X = x;
    B.ld = ld;
    return run_solver(ctx, nullptr, B, 1, 0, false, HD_SOLVER_PARTITION, 0, 0, nsys - 1, m);
}

int run_line_sweep(hd_ctx *ctx, const DevState *S, const DevState &s, int n, int dir, bool hyd, int start,
                   int end, double t, double s_start) {
    if (end < start) return HD_OK;
    const bool plain = ctx->cfg.form == HD_FORM_FP64_PLAIN;
    if (dir == 1)
        return plain ? sweep_impl<1, ROW_MAJOR>(ctx, S, s, n, hyd, start, end, t, s_start)
                     : sweep_impl<1, LINE_MAJOR>(ctx, S, s, n, hyd, start, end, t, s_start);
    return plain ? sweep_impl<0, ROW_MAJOR>(ctx, S, s, n, hyd, start, end, t, s_start)
                 : sweep_impl<0, LINE_MAJOR>(ctx, S, s, n, hyd, start, end, t, s_start);
}

}  // namespace hd
