// hd_kernels.hip - gfx950 kernels of the Hydist_C ADI timestep.
//
// Layout: every kernel takes `const DevState *S` (an array of scenario
// descriptors in device memory) and works on scenario blockIdx.z (2-D
// kernels) or blockIdx.y (line kernels), so one launch can step many
// scenarios of the same shape.
#include "hd_kernels.h"

namespace hd {

// ---------------------------------------------------------------------------
// Init: Onetime_init (support_funcs.cpp:21-69), per cell.
// ---------------------------------------------------------------------------
__global__ void k_onetime_init(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int M = s.M, N = s.N, W = s.P;
    if (i > N + 3 || j > M + 2) return;
    const double *h = s.h, *hsn = s.hsnham;
    int kh = 2;
    if (i >= 1 && i <= N + 1 && j >= 1 && j <= M + 1) {
        double sum = (h[(i - 1) * W + j - 1] + h[(i - 1) * W + j] + h[i * W + j - 1] + h[i * W + j]);
        if (sum * 0.25 > s.c.NANGDAY) {
            kh = 0;
            s.H_moi[i * W + j] = 0;
        }
        if (i <= N && j <= M) {
            const double ht = (h[(i - 1) * W + j - 1] + h[(i - 1) * W + j] + h[i * W + j - 1] + h[i * W + j]) * 0.25;
            s.htaiz[i * W + j] = ht;
            s.htaiz_bd[i * W + j] = ht;
            const double g = s.c.g, mu = s.c.mu_mn;
            if (h[i * W + j - 1] + h[i * W + j] != 0)
                s.Kx1[i * W + j] = g * pow((h[i * W + j - 1] + h[i * W + j]) * 0.5, -2 * mu) *
                                   pow((hsn[i * W + j - 1] + hsn[i * W + j]) * 0.5, 2.0);
            if (h[(i - 1) * W + j] + h[i * W + j] != 0)
                s.Ky1[i * W + j] = g * pow((h[(i - 1) * W + j] + h[i * W + j]) * 0.5, -2 * mu) *
                                   pow((hsn[(i - 1) * W + j] + hsn[i * W + j]) * 0.5, 2.0);
        }
    }
    s.khouot[i * W + j] = kh;
    s.segidH[i * W + j] = -1;
    s.segidV[i * W + j] = -1;
    s.segflV[2 * (i * W + j)] = -1;
    s.segflV[2 * (i * W + j) + 1] = -1;
}

// ---------------------------------------------------------------------------
// Segmentation (Find_Calculation_limits_Horizontal/Vertical,
// support_funcs.cpp:238-343) as run extraction: a run of wet cells clipped to
// [2, end] is a segment unless it starts at end-1 or later.  No cap on the
// number of segments per line.
// ---------------------------------------------------------------------------

// Rows (lines along j, unit stride): one wavefront per row; lane l takes
// position 2 + 64c + l of chunk c.  The khouot values of SEG_TILE chunks
// (and of the position after them) are loaded at once, unconditionally
// (clamped to the row), so one memory latency covers a whole tile; each
// chunk's wet ballot then gives its run starts (wet, previous position dry
// or p == 2) and ends (wet, next position dry or p == end) by bit shifts,
// and the lanes write them in order through a popcount prefix.  Block b
// takes rows 2 + b * SEG_ROWS_PER_BLOCK onwards, one per wavefront; its
// other wavefronts have nothing to do (few rows per block spread the row
// walks over all CUs).
constexpr int SEG_TILE = 16;

HD_DEV void segment_rows_block(const DevState &s, int b) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = b * SEG_ROWS_PER_BLOCK + wave + 2;
    if (wave >= SEG_ROWS_PER_BLOCK || i > s.N) return;
    const int end = s.M, W = s.P;
    const int *__restrict__ kh = s.khouot + i * W;
    int *first = s.segH_f + (size_t)i * s.capH, *last = s.segH_l + (size_t)i * s.capH;
    int nstart = 0, nend = 0, last_start = -1;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int nch = (end - 2) / 64 + 1;  // chunks covering positions 2..end
    uint64_t prevbit = 0;                 // wet bit of the position before the chunk (none before p = 2)
    for (int c0 = 0; c0 < nch; c0 += SEG_TILE) {
        int v[SEG_TILE];
#pragma unroll
        for (int u = 0; u < SEG_TILE; u++) v[u] = kh[min(2 + (c0 + u) * 64 + lane, end)];
        const int pn = 2 + (c0 + SEG_TILE) * 64;  // first position after the tile
        const int vn = kh[min(pn, end)];
        uint64_t w[SEG_TILE];
#pragma unroll
        for (int u = 0; u < SEG_TILE; u++) w[u] = __ballot(2 + (c0 + u) * 64 + lane <= end && v[u] == 0);
        const uint64_t wn = (pn <= end && vn == 0) ? 1ull : 0ull;
#pragma unroll
        for (int u = 0; u < SEG_TILE; u++) {
            if (c0 + u >= nch) break;
            const int base = 2 + (c0 + u) * 64, p = base + lane;
            const uint64_t nxt = u + 1 < SEG_TILE ? (w[u + 1] & 1ull) : wn;
            const uint64_t bs = w[u] & ~((w[u] << 1) | prevbit);
            const uint64_t be = w[u] & ~((w[u] >> 1) | (nxt << 63));
            prevbit = w[u] >> 63;
            if ((bs >> lane) & 1ull) first[nstart + __popcll(bs & below)] = p;
            if ((be >> lane) & 1ull) last[nend + __popcll(be & below)] = p;
            if (p <= end) {
                // segment of p: the starts at or before p; its run starts at the
                // last of them, and a run starting at end-1 or later is no segment
                const uint64_t le = bs & (below | (1ull << lane));
                const int start = le ? base + 63 - __clzll(le) : last_start;
                const bool in = ((w[u] >> lane) & 1ull) && start < end - 1;
                s.segidH[i * W + p] = in ? nstart + __popcll(le) - 1 : -1;
            }
            nstart += __popcll(bs);
            nend += __popcll(be);
            if (bs) last_start = base + 63 - __clzll(bs);
        }
    }
    // a run starting at end-1 or end is not a segment; it is always the last run
    int n = nstart;
    if (n > 0 && last_start >= end - 1) n--;
    if (lane == 0) s.segH_n[i] = n;
}

// Columns (lines along i, stride P): a workgroup takes SEG_COLS adjacent
// columns (64 bytes of khouot per row) and splits the rows 2..N into
// SEG_GROUPS contiguous groups, one thread per (group, column).  Pass 1
// counts run starts per group and finds its first dry row; LDS then gives
// each group its first segment index, the start of a run carried in from
// the groups before and the first dry row after it.  Pass 2 writes the
// run starts and ends, and per cell the segment index and first row;
// pass 3, backward, per cell the segment's last row.  Every pass loads
// SEG_TILE rows at once (clamped to the column), so one memory latency
// covers a batch of rows.
// Needs blockDim.x == SEG_COLS * SEG_GROUPS (every thread reaches the barrier).
HD_DEV void segment_cols_block(const DevState &s, int b) {
    __shared__ int cnt[SEG_GROUPS][SEG_COLS];
    __shared__ int lastst[SEG_GROUPS][SEG_COLS];
    __shared__ int firstdry[SEG_GROUPS][SEG_COLS];
    const int tx = threadIdx.x % SEG_COLS, g = threadIdx.x / SEG_COLS;
    const int j = b * SEG_COLS + tx + 2;
    const int end = s.N, W = s.P;
    const int rows = end - 1;  // rows 2..end
    const int per = (rows + SEG_GROUPS - 1) / SEG_GROUPS;
    const int r0 = 2 + g * per, r1 = min(end + 1, r0 + per);  // [r0, r1)
    const bool live = j <= s.M;
    const int *__restrict__ kh = s.khouot;
    int n = 0, ls = -1, fd = end + 1;
    if (live && r0 < r1) {
        bool prev = r0 > 2 && kh[(r0 - 1) * W + j] == 0;
        for (int p0 = r0; p0 < r1; p0 += SEG_TILE) {
            int v[SEG_TILE];
#pragma unroll
            for (int u = 0; u < SEG_TILE; u++) v[u] = kh[min(p0 + u, r1 - 1) * W + j];
#pragma unroll
            for (int u = 0; u < SEG_TILE; u++) {
                const int p = p0 + u;
                if (p < r1) {
                    const bool wet = v[u] == 0;
                    if (wet && !prev) {
                        n++;
                        ls = p;
                    }
                    if (!wet && fd > end) fd = p;
                    prev = wet;
                }
            }
        }
    }
    cnt[g][tx] = n;
    lastst[g][tx] = ls;
    firstdry[g][tx] = fd;
    __syncthreads();
    int off = 0, cs = -1;  // segments before the group; the last run start before it
    for (int q = 0; q < g; q++) {
        off += cnt[q][tx];
        cs = max(cs, lastst[q][tx]);
    }
    int ce = end + 1;  // the first dry row after the group (end + 1: none)
    for (int q = g + 1; q < SEG_GROUPS; q++) ce = min(ce, firstdry[q][tx]);
    // the column's last run start: a run starting at end-1 or later is no
    // segment, and it is always the last run
    int lst_all = -1;
    for (int q = 0; q < SEG_GROUPS; q++) lst_all = max(lst_all, lastst[q][tx]);
    const int dropped_from = lst_all >= end - 1 ? lst_all : end + 1;
    // Pass 2, forward: the tables, each cell's segment index and its
    // segment's first row (the run start, carried in from the groups
    // before when the run began there).  Pass 3, backward: each cell's
    // segment's last row (the row before the next dry one, carried in from
    // the groups after).  Cells outside segments get -1.
    int *fl = s.segflV;
    if (live && r0 < r1) {
        int *first = s.segV_f + (size_t)j * s.capV, *last = s.segV_l + (size_t)j * s.capV;
        int *segid = s.segidV + j;
        int k = off;  // starts at rows <= p
        bool prev = r0 > 2 && kh[(r0 - 1) * W + j] == 0;
        int start = cs;  // start of the run through p
        for (int p0 = r0; p0 < r1; p0 += SEG_TILE) {
            int v[SEG_TILE + 1];  // rows p0 .. p0 + SEG_TILE (the last for the end test)
#pragma unroll
            for (int u = 0; u <= SEG_TILE; u++) v[u] = kh[min(p0 + u, end) * W + j];
#pragma unroll
            for (int u = 0; u < SEG_TILE; u++) {
                const int p = p0 + u;
                if (p < r1) {
                    const bool wet = v[u] == 0;
                    const bool nwet = p < end && v[u + 1] == 0;
                    if (wet && !prev) {
                        first[k++] = p;
                        start = p;
                    }
                    if (wet && !nwet) last[k - 1] = p;
                    const bool in = wet && p < dropped_from;
                    segid[(size_t)p * W] = in ? k - 1 : -1;
                    fl[2 * ((size_t)p * W + j)] = in ? start : -1;
                    prev = wet;
                }
            }
        }
        int nd = ce;  // the first dry row after p
        for (int p0 = r0 + (r1 - 1 - r0) / SEG_TILE * SEG_TILE; p0 >= r0; p0 -= SEG_TILE) {
            int v[SEG_TILE];
#pragma unroll
            for (int u = 0; u < SEG_TILE; u++) v[u] = kh[min(p0 + u, r1 - 1) * W + j];
#pragma unroll
            for (int u = SEG_TILE - 1; u >= 0; u--) {
                const int p = p0 + u;
                if (p < r1) {
                    const bool wet = v[u] == 0;
                    if (!wet) nd = p;
                    fl[2 * ((size_t)p * W + j) + 1] = (wet && p < dropped_from) ? nd - 1 : -1;
                }
            }
        }
    }
    if (g == 0 && live) {
        int total = 0, lst = -1;
        for (int q = 0; q < SEG_GROUPS; q++) {
            total += cnt[q][tx];
            lst = max(lst, lastst[q][tx]);
        }
        // a run starting at end-1 or end is not a segment; it is always the last run
        if (total > 0 && lst >= end - 1) total--;
        s.segV_n[j] = total;
    }
}

// ---------------------------------------------------------------------------
// Htuongdoi (support_funcs.cpp:346-362) alone (init).
// ---------------------------------------------------------------------------
__global__ void k_htuongdoi(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 1 || j < 1 || i > s.N || j > s.M) return;
    const int W = s.P;
    const double *h = s.h, *z = s.z;
    s.Htdu[i * W + j] = (h[i * W + j - 1] + h[i * W + j] + z[(i + 1) * W + j] + z[i * W + j]) * 0.5;
    s.Htdv[i * W + j] = (h[(i - 1) * W + j] + h[i * W + j] + z[i * W + j + 1] + z[i * W + j]) * 0.5;
}

// ---------------------------------------------------------------------------
// preprocess_data (support_funcs.cpp:482-601): Q-boundary weights and
// hmax_*.  One workgroup; one thread per boundary segment sums its
// h^(5/3) terms in the reference's order.
// ---------------------------------------------------------------------------
__global__ void k_preprocess(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.x];
    const int M = s.M, N = s.N, W = s.P;
    const double ND = s.c.NANGDAY, dX = s.c.dX, ros = s.c.ros;
    const double *h = s.h;
    for (int side = 0; side < 4; side++) {
        const int *nseg, *fs, *ls;
        double *hs;
        if (side == 0) {
            nseg = s.segV_n + M; fs = s.segV_f + (size_t)M * s.capV; ls = s.segV_l + (size_t)M * s.capV;
            hs = s.hi;
        } else if (side == 1) {
            nseg = s.segV_n + 2; fs = s.segV_f + (size_t)2 * s.capV; ls = s.segV_l + (size_t)2 * s.capV;
            hs = s.hi + (N + 3);
        } else if (side == 2) {
            nseg = s.segH_n + 2; fs = s.segH_f + (size_t)2 * s.capH; ls = s.segH_l + (size_t)2 * s.capH;
            hs = s.hi + 2 * (N + 3);
        } else {
            nseg = s.segH_n + N; fs = s.segH_f + (size_t)N * s.capH; ls = s.segH_l + (size_t)N * s.capH;
            hs = s.hi + 2 * (N + 3) + (M + 3);
        }
        for (int k = threadIdx.x; k < *nseg; k += blockDim.x) {
            double sum = 0;
            for (int i = fs[k]; i <= ls[k]; i++) {
                double d;
                if (side == 0) d = h[i * W + M] + h[(i - 1) * W + M];
                else if (side == 1) d = h[i * W + 1] + h[(i - 1) * W + 1];
                else if (side == 2) d = h[1 * W + i] + h[1 * W + i - 1];
                else d = h[N * W + i] + h[N * W + i - 1];
                hs[i] = d / 2.0 - ND;
                sum += pow(hs[i], 5.0 / 3.0);
            }
            for (int i = fs[k]; i <= ls[k]; i++) hs[i] = pow(hs[i], 2.0 / 3.0) / (sum * dX);
        }
    }
    if (threadIdx.x == 0) {
        const double *hb = s.htaiz_bd;
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
        s.hmax[0] = hu;
        s.hmax[1] = hd;
        s.hmax[2] = hl;
        s.hmax[3] = hr;
    }
}

// ---------------------------------------------------------------------------
// Boundary update (Update_Boundary_Value, support_funcs.cpp:682-728).
// One workgroup per scenario; sides in the order up, down, left, right with
// a barrier between them so a later side wins at shared corner cells.
// Only segment 0 of the boundary line receives a value (:614-618).
// ---------------------------------------------------------------------------
// t_in is the half-step's time, or (from_clock, graph replays) the device
// clock holds it.
__global__ void __launch_bounds__(1024) k_boundary(const DevState *__restrict__ S, double t_in, int from_clock) {
    const DevState &s = S[blockIdx.x];
    const int M = s.M, N = s.N, W = s.P;
    const double t = (double)(float)(from_clock ? s.clock[0] : t_in);  // engine.cpp:22 passes t as float
    const int t1 = (int)(t / 3600);
    const double t2 = (t - (3600.0f * t1)) / 3600.0f;
    const double t2fs = (t - (3600.0 * t1)) / 3600.0;
    // series index clamped to the series (a time before 0 reads its first sample)
    const int k0 = max(0, min(t1, s.total_time - 1)), k1 = max(0, min(t1 + 1, s.total_time - 1));
    const int imax = max(M, N);
    // every side's segment-0 bounds and interpolated values first: nothing
    // below writes them, so their loads overlap instead of opening each side
    int n[4], f0[4], l0[4];
    double value[4], fsval[4];
#pragma unroll
    for (int side = 0; side < 4; side++) {
        const int *sn = side < 2 ? s.segV_n : s.segH_n;
        const int *sf = side < 2 ? s.segV_f : s.segH_f;
        const int *sl = side < 2 ? s.segV_l : s.segH_l;
        const size_t cap = side < 2 ? s.capV : s.capH;
        const int line = side == 0 ? M : (side == 3 ? N : 2);
        n[side] = sn[line];
        f0[side] = sf[(size_t)line * cap];
        l0[side] = sl[(size_t)line * cap];
        const double *bc = s.bc[side];
        value[side] = bc[k0] * (1.0f - t2) + bc[k1] * t2;
        fsval[side] = s.hasCC[side] ? s.CC[side][k0] * (1.0 - t2fs) + s.CC[side][k1] * t2fs : 0.0;
    }
    for (int side = 0; side < 4; side++) {
        const double *hs;
        double *barr;
        bool isU;
        int loc, ext;
        if (side == 0) {
            hs = s.hi; barr = s.vbt; isU = false; loc = M; ext = M + 1;
        } else if (side == 1) {
            hs = s.hi + (N + 3); barr = s.vbd; isU = false; loc = 2; ext = 1;
        } else if (side == 2) {
            hs = s.hi + 2 * (N + 3); barr = s.ubt; isU = true; loc = 2; ext = 1;
        } else {
            hs = s.hi + 2 * (N + 3) + (M + 3); barr = s.ubp; isU = true; loc = N; ext = N + 1;
        }
        if (n[side] > 0) {
            const double v = value[side], fv = fsval[side];
            const int a = f0[side], b = l0[side];
            const bool isQ = s.btype[side] == 2;
            for (int i = 2 + threadIdx.x; i <= imax; i += blockDim.x) {
                if (!(a <= i && i <= b)) continue;
                if (isQ) {
                    barr[i] = v * hs[i];
                    const int fi = isU ? loc : i, fj = isU ? i : loc;
                    s.FS[fi * W + fj] = fv * s.htaiz_bd[fi * W + fj] * s.hmax[side];
                } else if (isU) {
                    s.t_z[loc * W + i] = v;
                    s.t_z[ext * W + i] = v;
                } else {
                    s.t_z[i * W + loc] = v;
                    s.t_z[i * W + ext] = v;
                }
            }
            if (isQ && threadIdx.x == 0) {
                // prescribed discharge volume of this half-step (up/right sides face outward)
                const double sign = (side == 0 || side == 3) ? -1.0 : 1.0;
                s.acc[0] += sign * v * (0.5 * s.c.dT);
            }
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// Smoothing (Normalize, support_funcs.cpp:761-808): tmp = t_vel everywhere,
// Jacobi-smoothed on wet-interior cells [2,N] x [2,M].  The copy back into
// t_vel (update_buffer) is done by the next kernel that touches each cell.
// ---------------------------------------------------------------------------
template <bool ISU>
__global__ void k_smooth(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > s.N + 2 || j > s.M + 2) return;
    const int W = s.P;
    const double *__restrict__ val = ISU ? s.t_u : s.t_v;
    const int closenb = ISU ? W : 1;
    const int farnb = ISU ? 1 : W;
    const int pos = i * W + j;
    double out = val[pos];
    if (i >= 2 && j >= 2 && i <= s.N && j <= s.M && out != 0) {
        int count = 2;
        double nb = val[pos - closenb] + val[pos + closenb];
        if (s.khouot[pos - farnb] == 0) {
            nb += val[pos - farnb];
            count++;
        }
        if (s.khouot[pos + farnb] == 0) {
            nb += val[pos + farnb];
            count++;
        }
        out = out * s.c.heso + (1 - s.c.heso) * nb / count;
    }
    s.tmp[pos] = out;
}
template __global__ void k_smooth<true>(const DevState *);
template __global__ void k_smooth<false>(const DevState *);

HD_DEV bool interior(const DevState &s, int i, int j) { return i >= 2 && j >= 2 && i <= s.N && j <= s.M; }

// Whether position p of a line (khouot values `kh` at stride `st`) lies
// inside a segment and p+1 in the same segment: the range [first, last-1]
// that solveV / solveU cover (UVZSolver_multithread.cpp:955,974).
HD_DEV bool in_seg_not_last(const int *__restrict__ kh, int st, int p, int end) {
    if (p < 2 || p > end - 1) return false;
    if (kh[p * st] != 0 || kh[(p + 1) * st] != 0) return false;
    // the run through p must start at end-2 or earlier
    if (p <= end - 2) return true;
    return kh[(end - 2) * st] == 0;  // p == end-1
}

// solveV (UVZSolver_multithread.cpp:961-978) on every row segment cell
// j in [first, last-1], plus update_buffer of the smoothed t_u.
__global__ void k_solve_v(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (!interior(s, i, j)) return;
    const int W = s.P, pos = i * W + j;
    const double *tu_s = s.tmp;  // smoothed t_u
    if (in_seg_not_last(s.khouot + i * W, 1, j, s.M)) s.t_v[pos] = v_solver(s, tu_s, i, j);
    s.t_u[pos] = tu_s[pos];
}

// solveU (UVZSolver_multithread.cpp:938-959) on every column segment cell
// i in [first, last-1], plus update_buffer of the smoothed t_v.
__global__ void k_solve_u(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (!interior(s, i, j)) return;
    const int W = s.P, pos = i * W + j;
    const double *tv_s = s.tmp;  // smoothed t_v
    if (in_seg_not_last(s.khouot + j, W, i, s.N)) s.t_u[pos] = u_solver(s, tv_s, i, j);
    s.t_v[pos] = tv_s[pos];
}

// update_margin_elem_V / _U (UVZSolver_multithread.cpp:980-1060): one
// thread per line.  The reference walks the segments in order; only two of
// its writes can read the velocity: the first segment's start (a later
// segment starts past row/column 2, so its wall test holds and it writes
// 0) and the last segment's end (an earlier segment ends before `end`,
// so it writes 0).  Those two run first and last, and the zeros in
// between, at distinct cells (segments are separated by dry cells), are
// written from segment bounds loaded eight at a time, so the walk is not
// a chain of memory latencies.  The results are those of the ordered walk.
template <int DIR>
__global__ void k_margin(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.y];
    const int line = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int end = DIR == 1 ? s.M : s.N;  // lines: columns (U) or rows (V)
    if (line > end) return;
    const int W = s.P;
    const double ND = s.c.NANGDAY;
    const double *h = s.h;
    const int n = DIR == 1 ? s.segV_n[line] : s.segH_n[line];
    if (n <= 0) return;
    const int cap = DIR == 1 ? s.capV : s.capH;
    const int *fs = (DIR == 1 ? s.segV_f : s.segH_f) + (size_t)line * cap;
    const int *ls = (DIR == 1 ? s.segV_l : s.segH_l) + (size_t)line * cap;
    double *tv = DIR == 1 ? s.t_u : s.t_v;
    // cell q of the line
    auto at = [&](int q) { return DIR == 1 ? (size_t)q * W + line : (size_t)line * W + q; };
    {
        const int first = fs[0];
        bool ran1;
        if (DIR == 1) {
            const int j = line;
            ran1 = (first > 2) || ((first == 2) && ((fabs(h[1 * W + j] + h[1 * W + j - 1]) * 0.5 - ND) < 1e-16));
        } else {
            const int i = line;
            ran1 = (first > 2) ||
                   ((first == 2) && (fabs((h[i * W + first - 1] + h[(i - 1) * W + first - 1]) * 0.5 - ND) < 1e-16));
        }
        if (ran1)
            tv[at(first - 1)] = 0;
        else
            tv[at(first - 1)] = 2 * tv[at(first)] - tv[at(first + 1)];
    }
    for (int k0 = 0; k0 < n; k0 += 8) {
        int f[8], l[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = min(k0 + u, n - 1);
            f[u] = fs[k];
            l[u] = ls[k];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + u;
            if (k > 0 && k < n) tv[at(f[u] - 1)] = 0;  // start behind a wall
            if (k < n - 1) tv[at(l[u])] = 0;           // end before `end`
        }
    }
    {
        const int last = ls[n - 1];
        bool ran2;
        if (DIR == 1) {
            const int j = line, N = s.N;
            ran2 = (last < N) || ((last == N) && (fabs((h[N * W + j] + h[N * W + j - 1]) * 0.5 - ND) < 1e-16));
        } else {
            const int i = line, M = s.M;
            ran2 = (last < M) || ((last == M) && (fabs((h[i * W + last] + h[(i - 1) * W + last]) * 0.5 - ND) < 1e-16));
        }
        if (ran2)
            tv[at(last)] = 0;
        else
            tv[at(last)] = 2 * tv[at(last - 1)] - tv[at(last - 2)];
    }
}
template __global__ void k_margin<0>(const DevState *);
template __global__ void k_margin<1>(const DevState *);

// New water level at a cell after update_uvz: t_z inside [0,N]x[0,M],
// the old z elsewhere (update_uvz, support_funcs.cpp:750).
HD_DEV double z_after(const DevState &s, int i, int j) {
    const int pos = i * s.P + j;
    return (i <= s.N && j <= s.M) ? s.t_z[pos] : s.z[pos];
}

HD_DEV void vth_k_cell(const DevState &s, int i, int j, int pos, double htz, double H, double tu, double tu_w,
                       double tv, double tv_1, double htdu, double htdv, double h0, double h_1, double h_w);

// End of the first half-step (engine.cpp:122-139): update_buffer(v),
// update_h_moi, update_uvz and Htuongdoi in one pass; with vth, also the
// sediment pass's Find_VTH and hesoK (sediment_transport.cpp:539-602),
// which read only values this pass has at hand (the face velocities
// t_v at pos - 1 as update_buffer leaves them).
__global__ void k_post_a(const DevState *__restrict__ S, int vth) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > s.N + 2 || j > s.M + 2) return;
    const int W = s.P, pos = i * W + j;
    if (interior(s, i, j)) s.t_v[pos] = s.tmp[pos];
    const bool cell = i >= 1 && j >= 1 && i <= s.N && j <= s.M;
    const bool wet = cell && s.khouot[pos] == 0;
    if (wet) s.H_moi[pos] = s.htaiz[pos] + s.t_z[pos];
    const int kd = (int)s.c.kenhhepd, kn = (int)s.c.kenhhepng;
    if (i <= s.N && j <= s.M) {
        s.u[pos] = s.t_u[pos] * (1 - kd);
        s.v[pos] = (interior(s, i, j) ? s.tmp[pos] : s.t_v[pos]) * (1 - kn);
    }
    if (cell) {
        const double *h = s.h;
        const double zij = s.t_z[pos];
        const double htdu = (h[i * W + j - 1] + h[pos] + z_after(s, i + 1, j) + zij) * 0.5;
        const double htdv = (h[(i - 1) * W + j] + h[pos] + z_after(s, i, j + 1) + zij) * 0.5;
        s.Htdu[pos] = htdu;
        s.Htdv[pos] = htdv;
        if (vth) {
            const double H = wet ? s.htaiz[pos] + zij : s.H_moi[pos];
            const double tv = interior(s, i, j) ? s.tmp[pos] : s.t_v[pos];
            const double tv_1 = interior(s, i, j - 1) ? s.tmp[pos - 1] : s.t_v[pos - 1];
            vth_k_cell(s, i, j, pos, s.htaiz[pos], H, s.t_u[pos], s.t_u[pos - W], tv, tv_1, htdu, htdv, h[pos],
                       h[pos - 1], h[pos - W]);
        }
    }
    // z last: other threads read z only outside [0,N]x[0,M], which is not written
    if (i <= s.N && j <= s.M) s.z[pos] = s.t_z[pos];
}

// The half-step's time: set to `v`, or advanced by `v` (graph replays).
__global__ void k_clock(double *clock, double v, int advance) {
    if (threadIdx.x == 0) clock[0] = advance ? clock[0] + v : v;
}

// Second half-step before Reset_states_vertical (engine.cpp:229-236):
// update_buffer(u) and update_h_moi.  Also marks the rows at which the
// re-wetting scan can act: t_z > z with a dry (khouot == 1) neighbour in
// the column, from the values the scan starts with.
__global__ void k_pre_reset(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > s.N || j > s.M) return;
    const int W = s.P, pos = i * W + j;
    if (interior(s, i, j)) s.t_u[pos] = s.tmp[pos];
    if (i >= 1 && j >= 1 && s.khouot[pos] == 0) s.H_moi[pos] = s.htaiz[pos] + s.t_z[pos];
    if (i >= 2 && j >= 1 && s.t_z[pos] > s.z[pos] && (s.khouot[pos - W] == 1 || s.khouot[pos + W] == 1))
        atomicOr(&s.rwmask[(size_t)(i >> 6) * W + j], 1ull << (i & 63));
}

// Reset_states_vertical (support_funcs.cpp:142-188), re-wetting pass: one
// lane per column j in [1, M].  The reference scans i = 2..N in order; a
// row changes anything only if it is a marked row (see k_pre_reset) or the
// row below one whose scan re-wetted it, so the lane visits the marked rows
// in ascending order and follows each cascade downward while it lasts.
// Every update is written through, so later rows read current values.
HD_DEV bool rewet_row(const DevState &s, int i, int j) {
    const int W = s.P;
    double *t_z = s.t_z, *H_moi = s.H_moi, *FS = s.FS;
    int *kh = s.khouot;
    const double tzi = t_z[i * W + j];
    bool down = false;
    if (tzi > s.z[i * W + j]) {
        if (kh[(i - 1) * W + j] == 1) {
            t_z[(i - 1) * W + j] = tzi;
            H_moi[(i - 1) * W + j] = s.htaiz[(i - 1) * W + j] + tzi;
            kh[(i - 1) * W + j] = 0;
            FS[(i - 1) * W + j] = FS[i * W + j];
        }
        if (kh[(i + 1) * W + j] == 1) {
            t_z[(i + 1) * W + j] = tzi;
            H_moi[(i + 1) * W + j] = s.htaiz[(i + 1) * W + j] + tzi;
            kh[(i + 1) * W + j] = 0;
            FS[(i + 1) * W + j] = FS[i * W + j];
            down = true;
        }
    }
    return down;
}

__global__ void __launch_bounds__(64) k_rewet(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.y];
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > s.M) return;
    const int W = s.P, N = s.N, nw = (N >> 6) + 1;
    int next = 2;  // first row not yet scanned
    for (int w = 0; w < nw; w++) {
        unsigned long long m = s.rwmask[(size_t)w * W + j];
        if (!m) continue;
        s.rwmask[(size_t)w * W + j] = 0;
        while (m) {
            const int i = w * 64 + __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            if (i < next || i > N) continue;
            int r = i;
            // a row re-wetted from above is scanned next with its new level
            while (rewet_row(s, r, j) && r < N) r++;
            next = r + 1;
        }
    }
}

// Reset_states_vertical, drying pass: independent per cell once re-wetting
// is done (writes of t_u/t_v are zeros, so shared faces agree).
__global__ void k_dry(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 2 || j < 1 || i > s.N || j > s.M) return;
    const int W = s.P, pos = i * W + j;
    if ((s.khouot[pos] == 0) && (s.H_moi[pos] <= s.c.H_TINH)) {
        s.t_u[pos - W] = 0;
        s.t_u[pos] = 0;
        s.t_v[pos - 1] = 0;
        s.t_v[pos] = 0;
        s.khouot[pos] = 1;
        s.FS[pos] = 0;
    }
}

// End of the second half-step (engine.cpp:247-255): update_uvz and
// Htuongdoi; with vth, also Find_VTH and hesoK of the sediment pass that
// follows (the segmentation between them changes nothing they read).
__global__ void k_post_b(const DevState *__restrict__ S, int vth) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > s.N || j > s.M) return;
    const int W = s.P, pos = i * W + j;
    const int kd = (int)s.c.kenhhepd, kn = (int)s.c.kenhhepng;
    const double tu = s.t_u[pos], tv = s.t_v[pos];
    s.u[pos] = tu * (1 - kd);
    s.v[pos] = tv * (1 - kn);
    if (i >= 1 && j >= 1) {
        const double *h = s.h;
        const double zij = s.t_z[pos];
        const double htdu = (h[i * W + j - 1] + h[pos] + z_after(s, i + 1, j) + zij) * 0.5;
        const double htdv = (h[(i - 1) * W + j] + h[pos] + z_after(s, i, j + 1) + zij) * 0.5;
        s.Htdu[pos] = htdu;
        s.Htdv[pos] = htdv;
        if (vth)
            vth_k_cell(s, i, j, pos, s.htaiz[pos], s.H_moi[pos], tu, s.t_u[pos - W], tv, s.t_v[pos - 1], htdu, htdv,
                       h[pos], h[pos - 1], h[pos - W]);
    }
    s.z[pos] = s.t_z[pos];
}

// Interpolate_FS_ng on rows N and 2, then Interpolate_FS_d on column M
// (support_funcs.cpp:191-237, called from the segmentation kernels).
HD_DEV void interp_fs_block(const DevState &s) {
    const int M = s.M, N = s.N, W = s.P;
    const double dT = s.c.dT, dX = s.c.dX, dY = s.c.dY;
    double *FS = s.FS;
    for (int pass = 0; pass < 2; pass++) {
        const int location = pass == 0 ? N : 2, offset = pass == 0 ? N - 1 : 2, sign = pass == 0 ? -1 : 1;
        for (int j = 1 + threadIdx.x; j < M; j += blockDim.x) {
            if (s.t_u[location * W + j] > 0) {
                double deltaX = (s.u[location * W + j] + s.t_u[location * W + j]) * 0.25 * dT;
                // source rows outside the field, or not finite, are skipped;
                // tested before any conversion to int
                const double x = offset + sign * (deltaX / dX);
                if (!(x > -1.0 && x < N + 2.0)) continue;
                int i = (int)x;
                int i_deltaX = __double2int_rn(deltaX);
                int i_dX = __double2int_rn(dX);
                FS[location * W + j] =
                    FS[i * W + j] + sign * (FS[(i + 1) * W + j] - FS[i * W + j]) * (deltaX / dX - i_deltaX / i_dX);
            }
        }
        __syncthreads();
    }
    const int iend = min(M - 1, N + 2);
    const int location = M, offset = M - 1, sign = -1;
    for (int i = 1 + threadIdx.x; i <= iend; i += blockDim.x) {
        if (s.t_v[i * W + location] > 0) {
            double deltaY = (s.v[i * W + location] + s.t_v[i * W + location]) * 0.25 * dT;
            const double y = offset + sign * (deltaY / dY);
            if (!(y > -1.0 && y < M + 2.0)) continue;
            int j = (int)y;
            int j_deltaY = __double2int_rn(deltaY);
            int j_dY = __double2int_rn(dY);
            FS[i * W + location] = FS[i * W + j + 1] + (FS[i * W + j + 1] - FS[i * W + j]) * (deltaY / dY - j_deltaY / j_dY);
        }
    }
}

__global__ void __launch_bounds__(1024) k_interp_fs(const DevState *__restrict__ S) { interp_fs_block(S[blockIdx.x]); }

// Segmentation of both directions and the FS boundary interpolation in one
// launch (they touch disjoint data: khouot in, the segment tables out; FS
// on the outflow rows/column): blocks [0, row_blocks) take SEG_ROWS_PER_BLOCK
// rows each, the next col_blocks take SEG_COLS columns each, the last block
// the interpolation.  blockIdx.y is the scenario.
__global__ void __launch_bounds__(SEG_COLS *SEG_GROUPS) k_segment(const DevState *__restrict__ S, int row_blocks,
                                                                   int col_blocks) {
    const DevState &s = S[blockIdx.y];
    const int b = blockIdx.x;
    if (b < row_blocks)
        segment_rows_block(s, b);
    else if (b < row_blocks + col_blocks)
        segment_cols_block(s, b - row_blocks);
    else
        interp_fs_block(s);
}

// ---------------------------------------------------------------------------
// Sediment
// ---------------------------------------------------------------------------

// Find_VTH (sediment_transport.cpp:575-602) and hesoK (:539-573) of cell
// (i, j) in [1, N] x [1, M], from the values they read: htaiz, H_moi, t_u
// at pos and pos - W, t_v at pos and pos - 1, Htdu, Htdv, and h at pos,
// pos - 1 and pos - W.
HD_DEV void vth_k_cell(const DevState &s, int i, int j, int pos, double htz, double H, double tu, double tu_w,
                       double tv, double tv_1, double htdu, double htdv, double h0, double h_1, double h_w) {
    if (i >= 2 && j >= 2) {
        if ((htz > s.c.NANGDAY) && (H > s.c.H_TINH)) {
            double ut = (tu + tu_w) * 0.5;
            double vt = (tv + tv_1) * 0.5;
            s.VTH[pos] = sqrt(ut * ut + vt * vt);
        } else
            s.VTH[pos] = 0;
    }
    const double Ks = s.c.Ks, g = s.c.g;
    if (htdu > 0) {
        double Cz = 18 * log(12 * htdu / Ks);
        double k = 5.93 * sqrt(g) * (h_1 + h0) * 0.25 * fabs(tu) / Cz;
        s.Kx[pos] = fmin(100.0, fmax(5.0, k));
    }
    if (htdv > 0) {
        double Cz = 18 * log(12 * htdv / Ks);
        double k = 5.93 * sqrt(g) * (h_w + h0) * 0.25 * fabs(tv) / Cz;
        s.Ky[pos] = fmin(100.0, fmax(5.0, k));
    }
}

// The same as a pass of its own (the stage-by-stage entry points; a step
// forms them inside k_post_a / k_post_b).
__global__ void k_vth_k(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 1 || j < 1 || i > s.N || j > s.M) return;
    const int W = s.P, pos = i * W + j;
    vth_k_cell(s, i, j, pos, s.htaiz[pos], s.H_moi[pos], s.t_u[pos], s.t_u[pos - W], s.t_v[pos], s.t_v[pos - 1],
               s.Htdu[pos], s.Htdv[pos], s.h[pos], s.h[pos - 1], s.h[pos - W]);
}


// Update_FS (sediment_transport.cpp:654-663).
__global__ void k_update_fs(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 1 || j < 1 || i > s.N || j > s.M) return;
    const int pos = i * s.P + j;
    s.FS[pos] = s.tFS[pos];
}

// Calculate_Qb (sediment_transport.cpp:352-434), interior cells.  Q8: Qb is
// formed where VTH > 0 on wet cells and Toe is the coefficient.
__global__ void k_qb(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 2 || j < 2 || i > s.N || j > s.M) return;
    const int W = s.P, pos = i * W + j;
    const hd_coeffs &c = s.c;
    double qx = 0, qy = 0;
    double Toee = c.Toe, Tx = 0;
    const double vth = s.VTH[pos];
    if ((vth > 0) && (s.khouot[pos] == 0)) {
        double Tob = c.ro * s.Fw[pos] * (vth * vth);
        if (s.H_moi[pos] > c.ghtoe) Toee = c.Toe * (1 + c.hstoe * ((s.H_moi[pos] - c.ghtoe)));
        if (Tob > Toee) Tx = Tob / Toee - 1;
        // the constant powers come with the descriptor, computed once on
        // the host; the products run in the reference's order
        const double ptx = pow(Tx, 2.1);
        qx = 0.053 * s.qb_sg * s.qb_dm * ptx * s.Dxr_m03 * (s.t_u[pos] + s.t_u[pos - W]) * 0.5 / vth;
        qy = 0.053 * s.qb_sg * s.qb_dm * ptx * s.Dxr_m03 * (s.t_v[pos] + s.t_v[pos - 1]) * 0.5 / vth;
    }
    const int N = s.N, M = s.M;
    double *Qbx = s.Qbx, *Qby = s.Qby;
    Qbx[pos] = qx;
    Qby[pos] = qy;
    if (i == 2) { Qbx[pos - W] = qx; Qby[pos - W] = qy; }
    if (i == N) { Qbx[pos + W] = qx; Qby[pos + W] = qy; }
    if (j == 2) {
        Qbx[pos - 1] = qx; Qby[pos - 1] = qy;
        if (i == 2) { Qbx[1 * W + 1] = qx; Qby[1 * W + 1] = qy; }
        if (i == N) { Qbx[(N + 1) * W + 1] = qx; Qby[(N + 1) * W + 1] = qy; }
    }
    if (j == M) {
        Qbx[pos + 1] = qx; Qby[pos + 1] = qy;
        if (i == 2) { Qbx[1 * W + M + 1] = qx; Qby[1 * W + M + 1] = qy; }
        if (i == N) { Qbx[(N + 1) * W + M + 1] = qx; Qby[(N + 1) * W + M + 1] = qy; }
    }
}

// BedLoad -> _bed_load (sediment_transport.cpp:436-537, :664-673) on rows
// i in [3, N-2] and row segments of at least three cells.  Q9: the downwind
// test is t_u[pos] > 0.
//
// UPDATE_FS: Update_FS (FS = tFS on [1, N] x [1, M]) runs in the same pass;
// the bed load then reads the new FS from tFS, which holds the same values,
// so no thread reads a cell another thread is writing.
template <bool UPDATE_FS>
__global__ void k_bedload(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (UPDATE_FS && i >= 1 && j >= 1 && i <= s.N && j <= s.M) s.FS[i * s.P + j] = s.tFS[i * s.P + j];
    if (i < 3 || i > s.N - 2 || j < 2 || j > s.M) return;
    const int W = s.P, pos = i * W + j, M = s.M, N = s.N;
    const int k = s.segidH[pos];
    if (k < 0) return;
    const int first = s.segH_f[(size_t)i * s.capH + k], last = s.segH_l[(size_t)i * s.capH + k];
    if (last - first < 2) return;
    const hd_coeffs &c = s.c;
    const double dX = c.dX, dY = c.dY, dT = c.dT;
    const double *Qbx = s.Qbx, *Qby = s.Qby, *FS = UPDATE_FS ? s.tFS : s.FS;
    const int *kh = s.khouot;
    double p = 0, q = 0;
    if (kh[pos - W] == 1) {
        if (s.t_u[pos] < 0) p = (-3 * Qbx[pos] + 4 * Qbx[pos + W] - Qbx[pos + 2 * W]) / (dX * 2);
    } else {
        if (kh[pos + W] == 1) {
            if (s.t_u[pos] > 0) p = (3 * Qbx[pos] - 4 * Qbx[pos - W] + Qbx[pos - 2 * W]) / (dX * 2);
        } else
            p = (Qbx[pos + W] - Qbx[pos - W]) / (dX * 2);
    }
    if (kh[pos - 1] == 1) {
        if (s.t_v[pos] < 0) q = (-3 * Qby[pos] + 4 * Qby[pos + 1] - Qby[pos + 2]) / (dY * 2);
    } else {
        if (kh[pos + 1] == 1) {
            if (s.t_v[pos] > 0) q = (3 * Qby[pos] - 4 * Qby[pos - 1] + Qby[pos - 2]) / (dY * 2);
        } else
            q = (Qby[pos + 1] - Qby[pos - 1]) / (dY * 2);
    }
    double tH = p + q;
    p = FS[pos + W] - FS[pos];
    if ((kh[pos + 2 * W] == 1 || kh[pos - W] == 1) && (i < N)) p = 0;
    q = FS[pos] - FS[pos - W];
    if ((i > 2) && (kh[pos - 2 * W] == 1 || kh[pos + W] == 1)) q = 0;
    p = 1 / (dX * dX) * (s.Htdu[pos] * s.Kx[pos] * p - s.Htdu[pos - W] * s.Kx[pos - W] * q);
    tH = tH + p;
    p = FS[pos + 1] - FS[pos];
    if (((last < M) && (kh[pos - 1] == 1 || kh[pos + 2] == 1)) || (last == M && (kh[pos - 1] == 1))) p = 0;
    q = FS[pos] - FS[pos - 1];
    if (((first > 2) && (kh[pos - 2] == 1 || kh[pos + 1] == 1)) || (first == 2 && kh[pos + 1] == 1)) q = 0;
    p = 1 / ((dY * dY)) * (s.Htdv[pos] * s.Ky[pos] * p - s.Htdv[pos - 1] * s.Ky[pos - 1] * q);
    tH = tH + p;
    const double H = s.H_moi[pos], vth = s.VTH[pos], fsv = FS[pos];
    double cc = 18 * log(12 * H / c.Ks);
    double Uf = sqrt(c.g) * fabs(vth) / cc;
    double wsm = c.wss * pow4(1 - fsv);
    double Zf = 0, Sv = 0;
    if (fsv > 0.0) Zf = wsm / (0.4 * (Uf + 2 * wsm));
    if (s.ketdinh)
        Sv = source_chs(c, wsm, Zf, fsv, s.Fw[pos], vth, H);
    else
        Sv = source_nchs(c, wsm, Zf, Uf, H, fsv);
    s.dH[pos] += dT / (1 - c.Dorong) * (Sv + tH);
}
template __global__ void k_bedload<false>(const DevState *);
template __global__ void k_bedload<true>(const DevState *);

// Domain volume over wet compute cells, (htaiz + z) * dX * dY: per-row
// partial sums in column order, then rows in order (one workgroup).
__global__ void __launch_bounds__(1024) k_volume(const DevState *__restrict__ S) {
    const DevState &s = S[blockIdx.x];
    double *rowsum = s.acc + 8;
    const int W = s.P;
    for (int i = 2 + threadIdx.x; i <= s.N; i += blockDim.x) {
        double a = 0;
        for (int j = 2; j <= s.M; j++) {
            const int p = i * W + j;
            if (s.khouot[p] == 0) a += (s.htaiz[p] + s.z[p]) * s.c.dX * s.c.dY;
        }
        rowsum[i] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0;
        for (int i = 2; i <= s.N; i++) v += rowsum[i];
        s.acc[1] = v;
    }
}

// NaN/Inf watchdog: sets flag[0] = 1 when z, u or v of a wet compute cell
// is not finite.
__global__ void k_watchdog(const DevState *__restrict__ S, int *flag) {
    const DevState &s = S[blockIdx.z];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 2 || j < 2 || i > s.N || j > s.M) return;
    const int pos = i * s.P + j;
    if (s.khouot[pos] == 0 && !(isfinite(s.z[pos]) && isfinite(s.u[pos]) && isfinite(s.v[pos])))
        atomicOr(flag, 1);
}

// Sets a flag word (a kernel rather than a memset, so it can sit in a
// captured graph).
__global__ void k_set_flag(int *flag, int v) {
    if (threadIdx.x == 0) flag[0] = v;
}

}  // namespace hd
