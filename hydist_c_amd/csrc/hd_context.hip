// hd_context.hip - device allocation and upload of one scenario
// (attribute_arrays_memory_alloc / supporting_arrays_alloc,
// loader.cpp:224-375, with every buffer zero-initialised).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "hd_internal.h"

namespace hd {

namespace {

int round_up(int x, int m) { return (x + m - 1) / m * m; }

int upload2d(hd_ctx *ctx, double *dst, const double *src, const DevState &s) {
    if (!src) return HD_OK;
    const size_t W = s.M + 3, H = s.N + 3;
    // on the context's stream, after the buffer's zero-fill
    hipError_t e = hipMemcpy2DAsync(dst, s.P * sizeof(double), src, W * sizeof(double), W * sizeof(double), H,
                                    hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        ctx->err = std::string("upload: ") + hipGetErrorString(e);
        return HD_E_HIP;
    }
    return HD_OK;
}

}  // namespace

int alloc_scenario(hd_ctx *ctx, const hd_mesh *mesh, DevState &s) {
    s.M = mesh->M;
    s.N = mesh->N;
    s.P = ctx->cfg.dense_layout ? s.M + 3 : round_up(s.M + 3, 32);
    s.H = s.N + 4;
    s.capH = (s.M + 3) / 2 + 2;
    s.capV = (s.N + 3) / 2 + 2;
    s.total_time = mesh->total_time;
    for (int k = 0; k < 4; k++) {
        s.btype[k] = mesh->boundary_type[k];
        s.bienQ[k] = mesh->boundary_type[k] == 2 ? 1 : 0;
        s.hasCC[k] = mesh->CC[k] != nullptr;
    }
    s.quirks = ctx->cfg.quirks;
    s.ketdinh = ctx->ops.ketdinh ? 1 : 0;
    s.c = ctx->c;
    const size_t cells = (size_t)s.P * s.H;
    const size_t fbytes = cells * sizeof(double);
    int rc = 0;
    auto F = [&](double **p) {
        if (!rc) rc = ctx_alloc(ctx, (void **)p, fbytes);
    };
    auto A = [&](void **p, size_t bytes) {
        if (!rc) rc = ctx_alloc(ctx, p, bytes);
    };
    double *h, *hsn, *vis, *fw, *tx, *ty;
    F(&h); F(&hsn); F(&vis); F(&fw); F(&tx); F(&ty);
    F(&s.htaiz); F(&s.htaiz_bd); F(&s.Kx1); F(&s.Ky1);
    F(&s.u); F(&s.v); F(&s.z); F(&s.t_u); F(&s.t_v); F(&s.t_z); F(&s.H_moi); F(&s.Htdu); F(&s.Htdv);
    F(&s.FS); F(&s.tFS); F(&s.VTH); F(&s.Kx); F(&s.Ky); F(&s.Qbx); F(&s.Qby); F(&s.dH); F(&s.tmp);
    F(&s.keep0); F(&s.keep1);
    A((void **)&s.khouot, cells * sizeof(int));
    A((void **)&s.segidH, cells * sizeof(int));
    A((void **)&s.segidV, cells * sizeof(int));
    A((void **)&s.segflV, 2 * cells * sizeof(int));
    A((void **)&s.hi, 2 * (size_t)(s.M + s.N + 6) * sizeof(double));
    A((void **)&s.hmax, 4 * sizeof(double));
    A((void **)&s.ubt, (s.M + 3) * sizeof(double));
    A((void **)&s.ubp, (s.M + 3) * sizeof(double));
    A((void **)&s.vbt, (s.N + 3) * sizeof(double));
    A((void **)&s.vbd, (s.N + 3) * sizeof(double));
    for (int k = 0; k < 4; k++) {
        double *b = nullptr, *cc = nullptr;
        A((void **)&b, s.total_time * sizeof(double));
        A((void **)&cc, s.total_time * sizeof(double));
        s.bc[k] = b;
        s.CC[k] = cc;
    }
    A((void **)&s.segH_n, (s.N + 3) * sizeof(int));
    A((void **)&s.segH_f, (size_t)(s.N + 3) * s.capH * sizeof(int));
    A((void **)&s.segH_l, (size_t)(s.N + 3) * s.capH * sizeof(int));
    A((void **)&s.segV_n, (s.M + 3) * sizeof(int));
    A((void **)&s.segV_f, (size_t)(s.M + 3) * s.capV * sizeof(int));
    A((void **)&s.segV_l, (size_t)(s.M + 3) * s.capV * sizeof(int));
    A((void **)&s.acc, (8 + (size_t)s.H) * sizeof(double));
    A((void **)&s.rwmask, (size_t)((s.N >> 6) + 1) * s.P * sizeof(unsigned long long));
    if (rc) return rc;
    s.h = h; s.hsnham = hsn; s.visc = vis; s.Fw = fw; s.Tsxw = tx; s.Tsyw = ty;
    if ((rc = upload2d(ctx, h, mesh->h, s))) return rc;
    if ((rc = upload2d(ctx, hsn, mesh->hsnham, s))) return rc;
    if ((rc = upload2d(ctx, vis, mesh->VISCOIDX, s))) return rc;
    if ((rc = upload2d(ctx, fw, mesh->Fw, s))) return rc;
    if ((rc = upload2d(ctx, tx, mesh->Tsxw, s))) return rc;
    if ((rc = upload2d(ctx, ty, mesh->Tsyw, s))) return rc;
    if ((rc = upload2d(ctx, s.u, mesh->u0, s))) return rc;
    if ((rc = upload2d(ctx, s.v, mesh->v0, s))) return rc;
    if ((rc = upload2d(ctx, s.z, mesh->z0, s))) return rc;
    if ((rc = upload2d(ctx, s.FS, mesh->FS0, s))) return rc;
    for (int k = 0; k < 4; k++) {
        if (hipMemcpyAsync((void *)s.bc[k], mesh->bc[k], s.total_time * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream) != hipSuccess)
            return HD_E_HIP;
        if (mesh->CC[k] && hipMemcpyAsync((void *)s.CC[k], mesh->CC[k], s.total_time * sizeof(double),
                                          hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            return HD_E_HIP;
    }
    return HD_OK;
}

// Buffers shared by the scenarios of a context: the line-sweep buffers (one
// set per scenario, ROW_MAJOR [rho][line] and LINE_MAJOR [line][rho]
// layouts), the clock and the device descriptor array.
int alloc_shared(hd_ctx *ctx, int nscen) {
    const DevState &s = ctx->hs;
    int rc = 0;
    auto A = [&](void **p, size_t bytes) {
        if (!rc) rc = ctx_alloc(ctx, p, bytes);
    };
    SweepBufs &B = ctx->sweep;
    const int rowsU = 2 * (s.N - 1), rowsV = 2 * (s.M - 1);
    B.ld_rm[1] = round_up(s.M + 3, 64);
    B.ld_rm[0] = round_up(s.N + 3, 64);
    B.ld_lm[1] = round_up(rowsU, 64);
    B.ld_lm[0] = round_up(rowsV, 64);
    const size_t sw = std::max(std::max((size_t)rowsU * B.ld_rm[1], (size_t)rowsV * B.ld_rm[0]),
                               std::max((size_t)(s.M + 3) * B.ld_lm[1], (size_t)(s.N + 3) * B.ld_lm[0]));
    B.stride = (sw + 63) / 64 * 64;
    for (double **q : {&B.A, &B.B, &B.C, &B.D, &B.X, &B.Ap, &B.Bp}) A((void **)q, B.stride * nscen * sizeof(double));
    A((void **)&B.code, B.stride * nscen);
    const size_t nsys = std::max((size_t)s.capV * (s.M + 3), (size_t)s.capH * (s.N + 3));
    B.sstride = nsys;
    A((void **)&B.sys0, nsys * nscen * sizeof(int));
    A((void **)&B.sysn, nsys * nscen * sizeof(int));
    A((void **)&B.segl, nsys * nscen * sizeof(int));
    A((void **)&B.segcnt, std::max((size_t)64, nscen * sizeof(int)));
    A((void **)&ctx->clock, 64);
    A((void **)&ctx->wd_flag, 64);
    A((void **)&ctx->ds, sizeof(DevState) * nscen);
    if (rc) return rc;
    ctx->nscen = nscen;
    ctx->hs.clock = ctx->clock;
    for (DevState &d : ctx->hss) d.clock = ctx->clock;
    return sync_desc(ctx);
}


// Device bytes alloc_scenario + alloc_shared take for `nscen` scenarios of
// this mesh (the same formulas, so admission control can use it before
// allocating).
size_t footprint(const hd_mesh *mesh, int dense, int nscen) {
    const int M = mesh->M, N = mesh->N;
    const size_t P = dense ? M + 3 : round_up(M + 3, 32), H = N + 4;
    const size_t capH = (M + 3) / 2 + 2, capV = (N + 3) / 2 + 2;
    const size_t cells = P * H;
    auto al = [](size_t b) { return b ? b : (size_t)8; };
    size_t b = 30 * cells * 8 + 5 * cells * 4;
    b += al(2 * (size_t)(M + N + 6) * 8) + 32 + 2 * (M + 3) * 8 + 2 * (N + 3) * 8;
    b += 8 * al((size_t)mesh->total_time * 8);
    b += (N + 3) * 4 + 2 * (N + 3) * capH * 4 + (M + 3) * 4 + 2 * (M + 3) * capV * 4;
    b += (8 + H) * 8 + (size_t)((N >> 6) + 1) * P * 8;
    b *= nscen;
    const int rowsU = 2 * (N - 1), rowsV = 2 * (M - 1);
    const size_t sw = std::max(std::max((size_t)rowsU * round_up(M + 3, 64), (size_t)rowsV * round_up(N + 3, 64)),
                               std::max((size_t)(M + 3) * round_up(rowsU, 64), (size_t)(N + 3) * round_up(rowsV, 64)));
    const size_t stride = (sw + 63) / 64 * 64;
    const size_t nsys = std::max(capV * (M + 3), capH * (N + 3));
    b += 7 * stride * nscen * 8 + stride * nscen + 3 * nsys * nscen * 4 + std::max((size_t)64, nscen * sizeof(int)) +
         64 + 64 + sizeof(DevState) * nscen;
    return b;
}

}  // namespace hd
