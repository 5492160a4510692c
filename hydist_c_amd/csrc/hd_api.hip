// hd_api.hip - C ABI of the gfx950 Hydist_C timestep (include/hydist.h).
//
// A context owns one scenario's device buffers, its descriptor (DevState,
// kept both on the host and in device memory) and one HIP stream.  The time
// loop of Hydraulic_Calculation (engine.cpp:61-295) is a fixed sequence of
// kernel launches on that stream; nothing is synchronised until hd_sync or
// a copy-out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hd_internal.h"

using namespace hd;

namespace {

thread_local std::string g_err;

int set_err(hd_ctx *ctx, int code, const std::string &msg) {
    if (ctx)
        ctx->err = msg;
    else
        g_err = msg;
    return code;
}

#define HD_HIP(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_err((ctx), HD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace hd {

// Allocates and zero-fills on the context's stream, so the fill is ordered
// before every later upload and kernel on that stream (hipMemset would run
// on the null stream, which a non-blocking stream does not wait for).
int ctx_alloc(hd_ctx *ctx, void **p, size_t bytes) {
    void *q = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return set_err(ctx, HD_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(q, 0, bytes, ctx->stream);
    if (e != hipSuccess) return set_err(ctx, HD_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    ctx->allocs.push_back(q);
    ctx->bytes += bytes;
    *p = q;
    return HD_OK;
}

double *field_ptr(hd_ctx *ctx, int f) { return field_ptr(ctx->hs, f); }

double *field_ptr(const DevState &s, int f) {
    switch (f) {
        case HD_F_U: return s.u;
        case HD_F_V: return s.v;
        case HD_F_Z: return s.z;
        case HD_F_TU: return s.t_u;
        case HD_F_TV: return s.t_v;
        case HD_F_TZ: return s.t_z;
        case HD_F_HMOI: return s.H_moi;
        case HD_F_HTDU: return s.Htdu;
        case HD_F_HTDV: return s.Htdv;
        case HD_F_FS: return s.FS;
        case HD_F_TFS: return s.tFS;
        case HD_F_VTH: return s.VTH;
        case HD_F_KX: return s.Kx;
        case HD_F_KY: return s.Ky;
        case HD_F_QBX: return s.Qbx;
        case HD_F_QBY: return s.Qby;
        case HD_F_DH: return s.dH;
        case HD_F_HTAIZ: return s.htaiz;
        case HD_F_HTAIZ_BD: return s.htaiz_bd;
        case HD_F_KX1: return s.Kx1;
        case HD_F_KY1: return s.Ky1;
        case HD_F_H: return const_cast<double *>(s.h);
        case HD_F_HSNHAM: return const_cast<double *>(s.hsnham);
        case HD_F_VISCOIDX: return const_cast<double *>(s.visc);
        case HD_F_FW: return const_cast<double *>(s.Fw);
        default: return nullptr;
    }
}

Launch2D grid2d(const DevState &s, int nscen) {
    Launch2D l;
    l.block = dim3(64, 4, 1);
    l.grid = dim3((s.M + 3 + 63) / 64, (s.N + 4 + 3) / 4, nscen);
    return l;
}

// Pushes the host descriptor to the device copy.
// Pushes the host descriptors to the device copy.  Scenario 0 (hs) holds
// the parameters every scenario shares (coefficients, quirks, cohesion).
static void derive_constants(DevState &d) {
    d.r_dXbp = 1.0 / d.c.dXbp;
    d.r_dYbp = 1.0 / d.c.dYbp;
    d.r_dX2 = 1.0 / d.c.dX2;
    d.r_dY2 = 1.0 / d.c.dY2;
    d.Dxr_m03 = std::pow(d.c.Dxr, -0.3);
    d.qb_sg = std::pow((d.c.Sx - 1) * d.c.g, 0.5);
    d.qb_dm = std::pow(d.c.dm, 1.5);
}

int sync_desc(hd_ctx *ctx) {
    if (!ctx->ds) return HD_OK;
    derive_constants(ctx->hs);
    for (DevState &d : ctx->hss) {
        d.c = ctx->hs.c;
        d.quirks = ctx->hs.quirks;
        d.ketdinh = ctx->hs.ketdinh;
        d.clock = ctx->hs.clock;
        derive_constants(d);
    }
    HD_HIP(ctx, hipMemcpyAsync(ctx->ds, &ctx->hs, sizeof(DevState), hipMemcpyHostToDevice, ctx->stream));
    if (!ctx->hss.empty())
        HD_HIP(ctx, hipMemcpyAsync(ctx->ds + 1, ctx->hss.data(), ctx->hss.size() * sizeof(DevState),
                                   hipMemcpyHostToDevice, ctx->stream));
    return HD_OK;
}

// While a step is being captured into a graph (which replays with later
// times), the half-step advances the device clock by dT/2 and the kernels
// read the time from it; eager launches take t as an argument.
int set_clock(hd_ctx *ctx, double t) {
    (void)t;
    if (ctx->capturing)
        LAUNCH(ctx, "k_clock", k_clock<<<1, 64, 0, ctx->stream>>>(ctx->clock, 0.5 * ctx->c.dT, 1));
    return HD_OK;
}

int set_library_error(int code, const std::string &msg) { return set_err(nullptr, code, msg); }

int check_launch(hd_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(ctx, HD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return HD_OK;
}

// Whole-step graphs hold the launch shapes and kernels chosen when they were
// captured; every setting that changes either drops them, so the next
// graph step captures again.
void drop_graphs(hd_ctx *ctx) {
    for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}

static hipEvent_t pool_get(hd_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

hipEvent_t prof_begin(hd_ctx *ctx) {
    if (!ctx->prof) return nullptr;
    hipEvent_t e = pool_get(ctx);
    if (e) (void)hipEventRecord(e, ctx->stream);
    return e;
}

void prof_end(hd_ctx *ctx, const char *what, hipEvent_t e0) {
    if (!e0) return;
    hipEvent_t e1 = pool_get(ctx);
    if (!e1) return;
    (void)hipEventRecord(e1, ctx->stream);
    ctx->pending.push_back({what, e0, e1});
    if (ctx->pending.size() > 4096) prof_resolve(ctx);
}

int prof_resolve(hd_ctx *ctx) {
    for (auto &p : ctx->pending) {
        float ms = 0;
        HD_HIP(ctx, hipEventSynchronize(p.e1));
        HD_HIP(ctx, hipEventElapsedTime(&ms, p.e0, p.e1));
        auto &st = ctx->stage_stats[p.name];
        st.first += 1;
        st.second += ms;
        ctx->event_pool.push_back(p.e0);
        ctx->event_pool.push_back(p.e1);
    }
    ctx->pending.clear();
    return HD_OK;
}

// ---------------------------------------------------------------------------
// Stage sequences.  `S` is the device descriptor array, `n` the number of
// scenarios it holds (all of the same shape as `s`).
// ---------------------------------------------------------------------------
int run_segmentation(hd_ctx *ctx, const DevState *S, const DevState &s, int n) {
    hipStream_t st = ctx->stream;
    const int rb = (s.N - 1 + SEG_ROWS_PER_BLOCK - 1) / SEG_ROWS_PER_BLOCK;
    const int cb = (s.M - 1 + SEG_COLS - 1) / SEG_COLS;
    LAUNCH(ctx, "k_segment", k_segment<<<dim3(rb + cb + 1, n), SEG_COLS * SEG_GROUPS, 0, st>>>(S, rb, cb));
    return HD_OK;
}

int run_init(hd_ctx *ctx, const DevState *S, const DevState &s, int n) {
    hipStream_t st = ctx->stream;
    Launch2D g = grid2d(s, n);
    LAUNCH(ctx, "k_onetime_init", k_onetime_init<<<g.grid, g.block, 0, st>>>(S));
    int rc = run_segmentation(ctx, S, s, n);
    if (rc) return rc;
    LAUNCH(ctx, "k_htuongdoi", k_htuongdoi<<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_preprocess", k_preprocess<<<n, 64, 0, st>>>(S));
    return HD_OK;
}

int run_halfstep_uz(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t) {
    hipStream_t st = ctx->stream;
    const hd_options &ops = ctx->ops;
    Launch2D g = grid2d(s, n);
    int rc = set_clock(ctx, t);
    if (rc) return rc;
    if (!ops.channel) LAUNCH(ctx, "k_boundary", k_boundary<<<n, 1024, 0, st>>>(S, t, ctx->capturing ? 1 : 0));
    int start = 2, end = s.M;
    if (ops.channel && ops.kenhhepd) {
        start = 3;
        end = s.M - 1;
    }
    rc = run_sweep(ctx, S, s, n, 1, start, end);
    if (rc) return rc;
    LAUNCH(ctx, "k_smooth<u>", k_smooth<true><<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_solve_v", k_solve_v<<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_margin<v>", k_margin<0><<<dim3((s.N - 1 + 63) / 64, n), 64, 0, st>>>(S));
    LAUNCH(ctx, "k_smooth<v>", k_smooth<false><<<g.grid, g.block, 0, st>>>(S));
    // (with sediment, k_post_a also forms VTH and the diffusion
    // coefficients of the FS pass below)
    const bool sed = ctx->cfg.sediment && t >= ops.sediment_start;
    LAUNCH(ctx, sed ? "k_post_a+vth" : "k_post_a", k_post_a<<<g.grid, g.block, 0, st>>>(S, sed ? 1 : 0));
    // khouot is unchanged in this half-step, so the segment tables are too;
    // only the FS boundary interpolation of Find_Calculation_limits runs.
    LAUNCH(ctx, "k_interp_fs", k_interp_fs<<<n, 1024, 0, st>>>(S));
    if (sed) {
        rc = run_line_sweep(ctx, S, s, n, 1, false, 3, s.M - 1, t, ops.bed_change_start);
        if (rc) return rc;
        LAUNCH(ctx, "k_update_fs", k_update_fs<<<g.grid, g.block, 0, st>>>(S));
    }
    return HD_OK;
}

int run_halfstep_vz_a(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t) {
    hipStream_t st = ctx->stream;
    const hd_options &ops = ctx->ops;
    Launch2D g = grid2d(s, n);
    int rc = set_clock(ctx, t);
    if (rc) return rc;
    if (!ops.channel) LAUNCH(ctx, "k_boundary", k_boundary<<<n, 1024, 0, st>>>(S, t, ctx->capturing ? 1 : 0));
    int start = 2, end = s.N;
    if (ops.channel && ops.kenhhepng) start = 3;
    rc = run_sweep(ctx, S, s, n, 0, start, end);
    if (rc) return rc;
    LAUNCH(ctx, "k_smooth<v>", k_smooth<false><<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_solve_u", k_solve_u<<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_margin<u>", k_margin<1><<<dim3((s.M - 1 + 63) / 64, n), 64, 0, st>>>(S));
    LAUNCH(ctx, "k_smooth<u>", k_smooth<true><<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_pre_reset", k_pre_reset<<<g.grid, g.block, 0, st>>>(S));
    LAUNCH(ctx, "k_rewet", k_rewet<<<dim3((s.M + 63) / 64, n), 64, 0, st>>>(S));
    LAUNCH(ctx, "k_dry", k_dry<<<g.grid, g.block, 0, st>>>(S));
    return HD_OK;
}

// From update_uvz on (engine.cpp:247-288); save_result sits between the two
// parts in the reference's loop (engine.cpp:243-244).
int run_halfstep_vz_b(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t) {
    hipStream_t st = ctx->stream;
    const hd_options &ops = ctx->ops;
    Launch2D g = grid2d(s, n);
    const bool sed = ctx->cfg.sediment && t >= ops.sediment_start;
    LAUNCH(ctx, sed ? "k_post_b+vth" : "k_post_b", k_post_b<<<g.grid, g.block, 0, st>>>(S, sed ? 1 : 0));
    int rc = run_segmentation(ctx, S, s, n);
    if (rc) return rc;
    if (sed) {
        rc = run_line_sweep(ctx, S, s, n, 0, false, 3, s.N - 1, t, ops.bed_change_start);
        if (rc) return rc;
        if (ctx->cfg.bed_change && t >= ops.bed_change_start) {
            // Calculate_Qb does not read FS, so Update_FS can wait for (and
            // run inside) the BedLoad pass
            LAUNCH(ctx, "k_qb", k_qb<<<g.grid, g.block, 0, st>>>(S));
            LAUNCH(ctx, "k_bedload", k_bedload<true><<<g.grid, g.block, 0, st>>>(S));
        } else {
            LAUNCH(ctx, "k_update_fs", k_update_fs<<<g.grid, g.block, 0, st>>>(S));
        }
    }
    return HD_OK;
}

int run_halfstep_vz(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t) {
    int rc = run_halfstep_vz_a(ctx, S, s, n, t);
    if (rc) return rc;
    return run_halfstep_vz_b(ctx, S, s, n, t);
}

int run_sweep(hd_ctx *ctx, const DevState *S, const DevState &s, int n, int dir, int start, int end) {
    return run_line_sweep(ctx, S, s, n, dir, true, start, end, 0.0, 0.0);
}

int check_setup(const hd_coeffs *coeffs, const hd_config *cfg) {
    const int form = cfg ? cfg->form : HD_FORM_FP64_PLAIN;
    if (form < HD_FORM_FP64_PLAIN || form > HD_FORM_FP32_REFINE)
        return set_err(nullptr, HD_E_INVALID, "unknown form");
    if (cfg && (cfg->quirks & ~HD_QUIRK_Q11_LITERAL))
        return set_err(nullptr, HD_E_INVALID, "unknown quirk bits (only HD_QUIRK_Q11_LITERAL is defined)");
    for (double v : {coeffs->dX, coeffs->dY, coeffs->dT})
        if (!(v > 0.0) || !std::isfinite(v))  // NaN fails v > 0
            return set_err(nullptr, HD_E_INVALID, "dX, dY and dT must be positive and finite");
    return HD_OK;
}

}  // namespace hd

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int hd_finalize_coeffs(hd_coeffs *ce) {
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
    const double aw = std::fabs(ce->Wind);
    ce->Windx = 0.0013 * (0.00075 + 0.000067 * aw) * aw * ce->Wind * std::cos(ce->huonggio * (PI / 180));
    ce->Windy = 0.0013 * (0.00075 + 0.000067 * aw) * aw * ce->Wind * std::sin(ce->huonggio * (PI / 180));
    ce->Ufr = 0.25 * std::pow((ce->Sx - 1) * ce->g, 8.0 / 15.0) * std::pow(ce->dm, 9.0 / 15.0) *
              std::pow(ce->muy, -1.0 / 15.0);
    // loader.cpp:85 evaluates g*(Sx-1)/muy*muy left to right; kept as written
    ce->Dxr = ce->dm * std::pow(ce->g * (ce->Sx - 1) / ce->muy * ce->muy, (1.0 / 3.0));
    ce->wss = 10 * ce->muy / ce->dm *
              (std::sqrt(1 + 0.01 * (ce->Sx - 1) * 9.81 * std::pow(ce->dm, 3) / std::pow(ce->muy, 2)) - 1);
    return HD_OK;
}

int hd_load_coeffs(hd_coeffs *ce) {
    if (!ce) return HD_E_INVALID;
    std::memset(ce, 0, sizeof(*ce));
    // loader.cpp:6-78
    ce->dX = 10.0;
    ce->dY = 10.0;
    ce->dT = 2;
    ce->NANGDAY = 0.0;
    ce->H_TINH = 0.02;
    ce->heso = 0.94;
    ce->mu_mn = 0.2;
    ce->NDbphai = ce->NDbtrai = ce->NDbtren = ce->NDbduoi = 0.5;
    ce->Tod = 0.06;
    ce->Toe = 0.15;
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
    ce->g = 9.81;
    return hd_finalize_coeffs(ce);
}

// Reads the device descriptors back and compares them with the host copies,
// so a descriptor that did not arrive intact is an error before any kernel
// reads it.
static int verify_desc(hd_ctx *ctx) {
    std::vector<DevState> back(ctx->nscen);
    if (hipMemcpy(back.data(), ctx->ds, back.size() * sizeof(DevState), hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(ctx, HD_E_HIP, "descriptor read-back failed");
    for (int k = 0; k < ctx->nscen; k++) {
        const DevState &h = k == 0 ? ctx->hs : ctx->hss[k - 1];
        if (std::memcmp(&back[k], &h, sizeof(DevState)) != 0)
            return set_err(ctx, HD_E_HIP, "device descriptor " + std::to_string(k) + " does not match the host copy");
    }
    return HD_OK;
}

static int check_mesh(const hd_mesh *mesh) {
    if (mesh->M < 4 || mesh->N < 4 || !mesh->h || !mesh->hsnham || !mesh->VISCOIDX)
        return set_err(nullptr, HD_E_INVALID, "mesh needs M, N >= 4 and h, hsnham, VISCOIDX");
    if (mesh->total_time < 2) return set_err(nullptr, HD_E_INVALID, "boundary series need >= 2 samples");
    for (int s = 0; s < 4; s++)
        if (!mesh->bc[s]) return set_err(nullptr, HD_E_INVALID, "boundary series bc[side] is NULL");
    return HD_OK;
}

// One context for n scenarios of one shape (n = 1: hd_create).
static int create_ctx(const hd_mesh *meshes, int n, const hd_coeffs *coeffs, const hd_options *ops,
                      const hd_config *cfg, hd_ctx **out) {
    if (!out) return set_err(nullptr, HD_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!meshes || !coeffs || !ops || n < 1) return set_err(nullptr, HD_E_INVALID, "mesh/coeffs/ops is NULL");
    for (int k = 0; k < n; k++) {
        int rc = check_mesh(&meshes[k]);
        if (rc) return rc;
        if (meshes[k].M != meshes[0].M || meshes[k].N != meshes[0].N ||
            meshes[k].total_time != meshes[0].total_time)
            return set_err(nullptr, HD_E_INVALID, "batched scenarios need one M, N and total_time");
    }
    int rc0 = hd::check_setup(coeffs, cfg);
    if (rc0) return rc0;
    const int form = cfg ? cfg->form : HD_FORM_FP64_PLAIN;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_err(nullptr, HD_E_NODEV, "no HIP device");
    const int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= ndev) return set_err(nullptr, HD_E_NODEV, "device ordinal out of range");

    hd_ctx *ctx = new hd_ctx();
    ctx->device = dev;
    ctx->c = *coeffs;
    ctx->ops = *ops;
    ctx->ops.M = meshes[0].M;
    ctx->ops.N = meshes[0].N;
    if (cfg) ctx->cfg = *cfg;
    ctx->cfg.form = form;
    if (hipSetDevice(dev) != hipSuccess) {
        delete ctx;
        return set_err(nullptr, HD_E_NODEV, "hipSetDevice failed");
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return set_err(nullptr, HD_E_HIP, "hipStreamCreate failed");
    }
    int rc = hd::alloc_scenario(ctx, &meshes[0], ctx->hs);
    for (int k = 1; k < n && !rc; k++) {
        DevState d{};
        rc = hd::alloc_scenario(ctx, &meshes[k], d);
        if (!rc) ctx->hss.push_back(d);
    }
    if (!rc) rc = hd::alloc_shared(ctx, n);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = set_err(ctx, HD_E_HIP, "upload failed");
    if (!rc) rc = verify_desc(ctx);
    if (rc) {
        g_err = ctx->err;
        hd_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return HD_OK;
}

int hd_create(const hd_mesh *mesh, const hd_coeffs *coeffs, const hd_options *ops, const hd_config *cfg,
              hd_ctx **out) {
    return create_ctx(mesh, 1, coeffs, ops, cfg, out);
}

int hd_init(hd_ctx *ctx) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    HD_HIP(ctx, hipSetDevice(ctx->device));
    int rc = run_init(ctx, ctx->ds, ctx->hs, ctx->nscen);
    if (rc) return rc;
    ctx->inited = true;
    return HD_OK;
}

int hd_update_boundary(hd_ctx *ctx, double t) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    // the Q sides' weights hi come from hd_init (preprocess_data)
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    LAUNCH(ctx, "k_boundary", k_boundary<<<ctx->nscen, 1024, 0, ctx->stream>>>(ctx->ds, t, 0));
    return HD_OK;
}

int hd_halfstep_uz(hd_ctx *ctx, double t) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    if (ctx->fault_stage == 1) return set_err(ctx, HD_E_FAULT, "injected fault in UZ half-step");
    return run_halfstep_uz(ctx, ctx->ds, ctx->hs, ctx->nscen, t);
}

int hd_halfstep_vz(hd_ctx *ctx, double t) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    if (ctx->fault_stage == 2) return set_err(ctx, HD_E_FAULT, "injected fault in VZ half-step");
    return run_halfstep_vz(ctx, ctx->ds, ctx->hs, ctx->nscen, t);
}

// Which optional stages a step from t runs: sediment in the first and the
// second half-step, bed change in the second.
static int step_key(const hd_ctx *ctx, double t) {
    const double dT = ctx->c.dT, t1 = t + 0.5 * dT, t2 = t1 + dT * 0.5;
    const hd_options &o = ctx->ops;
    const bool sed = ctx->cfg.sediment != 0;
    return (sed && t1 >= o.sediment_start ? 1 : 0) | (sed && t2 >= o.sediment_start ? 2 : 0) |
           (sed && ctx->cfg.bed_change && t2 >= o.sediment_start && t2 >= o.bed_change_start ? 4 : 0);
}

static int step_eager(hd_ctx *ctx, double *t) {
    const double dT = ctx->c.dT;
    *t += 0.5 * dT;
    int rc = hd_halfstep_uz(ctx, *t);
    if (rc) return rc;
    *t += dT * 0.5;
    return hd_halfstep_vz(ctx, *t);
}

// A whole step as one graph launch.  The graph for a stage pattern is
// captured once; it advances the device clock itself, so a replay needs
// only the clock set to the step's start time.
static int step_graph(hd_ctx *ctx, double *t) {
    const int key = step_key(ctx, *t);
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        if (ctx->prof) {
            int rc = prof_resolve(ctx);
            if (rc) return rc;
        }
        const bool prof = ctx->prof;
        ctx->prof = false;
        HD_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        ctx->capturing = true;
        double tt = *t;
        int rc = step_eager(ctx, &tt);
        ctx->capturing = false;
        ctx->prof = prof;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_err(ctx, HD_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_err(ctx, HD_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        it = ctx->graphs.emplace(key, x).first;
    }
    LAUNCH(ctx, "k_clock", k_clock<<<1, 64, 0, ctx->stream>>>(ctx->clock, *t, 0));
    HD_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
    *t += ctx->c.dT;
    return HD_OK;
}

// Reads the watchdog's flag (the stream must be idle): HD_E_NUMERIC once a
// check has found a non-finite value.
static int watchdog_status(hd_ctx *ctx) {
    if (!ctx->wd_tripped && ctx->wd_every > 0) {
        int flag = 0;
        HD_HIP(ctx, hipMemcpy(&flag, ctx->wd_flag, sizeof(int), hipMemcpyDeviceToHost));
        ctx->wd_tripped = flag != 0;
    }
    if (ctx->wd_tripped)
        return set_err(ctx, HD_E_NUMERIC, "watchdog: non-finite z, u or v in a wet cell (checked every " +
                                              std::to_string(ctx->wd_every) + " steps)");
    return HD_OK;
}

// One full step, then the watchdog's check when one is due; *checked tells
// whether a check was queued.
static int step_checked(hd_ctx *ctx, double *t, bool *checked) {
    *checked = false;
    int rc = (ctx->use_graph && !ctx->fault_stage) ? step_graph(ctx, t) : step_eager(ctx, t);
    if (rc || ctx->wd_every <= 0) return rc;
    if (++ctx->wd_steps % ctx->wd_every == 0) {
        Launch2D g = grid2d(ctx->hs, ctx->nscen);
        LAUNCH(ctx, "k_watchdog", k_watchdog<<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds, ctx->wd_flag));
        *checked = true;
    }
    return HD_OK;
}

int hd_step(hd_ctx *ctx, double *t) {
    if (!ctx || !t) return set_err(ctx, HD_E_INVALID, "ctx or t is NULL");
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    bool checked;
    return step_checked(ctx, t, &checked);
}

// A step of hd_run / hd_run_cb: waits for the watchdog's check when one ran.
static int run_step(hd_ctx *ctx, double *t) {
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    bool checked;
    int rc = step_checked(ctx, t, &checked);
    if (rc || !checked) return rc;
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return watchdog_status(ctx);
}

int hd_run(hd_ctx *ctx, double *t, double Tmax) {
    if (!ctx || !t) return set_err(ctx, HD_E_INVALID, "ctx or t is NULL");
    if (ctx->wd_tripped) return watchdog_status(ctx);
    const int T = (int)Tmax;
    while (*t < T) {
        int rc = run_step(ctx, t);
        if (rc) return rc;
    }
    return HD_OK;
}

int hd_run_cb(hd_ctx *ctx, double *t, double Tmax, int interval, hd_output_cb cb, void *user) {
    if (!ctx || !t || interval < 1) return set_err(ctx, HD_E_INVALID, "ctx/t NULL or interval < 1");
    if (ctx->wd_tripped) return watchdog_status(ctx);
    const int T = (int)Tmax;
    while (*t < T) {
        int rc = run_step(ctx, t);
        if (rc) return rc;
        if (cb && ((int)*t % interval == 0) && (*t - (int)*t == 0)) {
            HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (cb(ctx, *t, user)) return set_err(ctx, HD_E_CANCELLED, "stopped by the output callback");
        }
    }
    return HD_OK;
}

int hd_set_watchdog(hd_ctx *ctx, int every) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (every < 0) return set_err(ctx, HD_E_INVALID, "watchdog interval must be >= 0");
    LAUNCH(ctx, "k_set_flag", k_set_flag<<<1, 64, 0, ctx->stream>>>(ctx->wd_flag, 0));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->wd_every = every;
    ctx->wd_steps = 0;
    ctx->wd_tripped = false;
    return HD_OK;
}

int hd_inject_fault(hd_ctx *ctx, int stage) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (stage < 0 || stage > 2) return set_err(ctx, HD_E_INVALID, "fault stage must be 0 (off), 1 (UZ) or 2 (VZ)");
    ctx->fault_stage = stage;
    return HD_OK;
}

static int stage_guard(hd_ctx *ctx) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    return HD_OK;
}

int hd_sediment_j(hd_ctx *ctx, double t) {
    int rc = stage_guard(ctx);
    if (rc) return rc;
    const DevState &s = ctx->hs;
    const int n = ctx->nscen;
    Launch2D g = grid2d(s, n);
    LAUNCH(ctx, "k_vth_k", k_vth_k<<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds));
    rc = run_line_sweep(ctx, ctx->ds, s, n, 1, false, 3, s.M - 1, t, ctx->ops.bed_change_start);
    if (rc) return rc;
    LAUNCH(ctx, "k_update_fs", k_update_fs<<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds));
    return HD_OK;
}

int hd_sediment_i(hd_ctx *ctx, double t) {
    int rc = stage_guard(ctx);
    if (rc) return rc;
    const DevState &s = ctx->hs;
    const int n = ctx->nscen;
    Launch2D g = grid2d(s, n);
    LAUNCH(ctx, "k_vth_k", k_vth_k<<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds));
    rc = run_line_sweep(ctx, ctx->ds, s, n, 0, false, 3, s.N - 1, t, ctx->ops.bed_change_start);
    if (rc) return rc;
    LAUNCH(ctx, "k_update_fs", k_update_fs<<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds));
    return HD_OK;
}

int hd_bedload(hd_ctx *ctx, double t) {
    (void)t;
    int rc = stage_guard(ctx);
    if (rc) return rc;
    Launch2D g = grid2d(ctx->hs, ctx->nscen);
    LAUNCH(ctx, "k_qb", k_qb<<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds));
    LAUNCH(ctx, "k_bedload", k_bedload<false><<<g.grid, g.block, 0, ctx->stream>>>(ctx->ds));
    return HD_OK;
}

int hd_sync(hd_ctx *ctx) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return watchdog_status(ctx);
}

int hd_copy_out(hd_ctx *ctx, int field, double *dst, size_t n) {
    if (!ctx || !dst) return set_err(ctx, HD_E_INVALID, "NULL argument");
    double *src = field_ptr(ctx, field);
    const size_t W = ctx->hs.M + 3, H = ctx->hs.N + 3;
    if (!src || n != W * H) return set_err(ctx, HD_E_INVALID, "bad field id or size");
    HD_HIP(ctx, hipMemcpy2DAsync(dst, W * sizeof(double), src, ctx->hs.P * sizeof(double), W * sizeof(double), H,
                                 hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_copy_in(hd_ctx *ctx, int field, const double *src, size_t n) {
    if (!ctx || !src) return set_err(ctx, HD_E_INVALID, "NULL argument");
    double *dst = field_ptr(ctx, field);
    const size_t W = ctx->hs.M + 3, H = ctx->hs.N + 3;
    if (!dst || n != W * H) return set_err(ctx, HD_E_INVALID, "bad field id or size");
    HD_HIP(ctx, hipMemcpy2DAsync(dst, ctx->hs.P * sizeof(double), src, W * sizeof(double), W * sizeof(double), H,
                                 hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_copy_out_i32(hd_ctx *ctx, int field, int32_t *dst, size_t n) {
    if (!ctx || !dst || field != HD_F_KHOUOT) return set_err(ctx, HD_E_INVALID, "bad argument");
    const size_t W = ctx->hs.M + 3, H = ctx->hs.N + 3;
    if (n != W * H) return set_err(ctx, HD_E_INVALID, "bad size");
    HD_HIP(ctx, hipMemcpy2DAsync(dst, W * sizeof(int32_t), ctx->hs.khouot, ctx->hs.P * sizeof(int32_t),
                                 W * sizeof(int32_t), H, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_copy_in_i32(hd_ctx *ctx, int field, const int32_t *src, size_t n) {
    if (!ctx || !src || field != HD_F_KHOUOT) return set_err(ctx, HD_E_INVALID, "bad argument");
    const size_t W = ctx->hs.M + 3, H = ctx->hs.N + 3;
    if (n != W * H) return set_err(ctx, HD_E_INVALID, "bad size");
    HD_HIP(ctx, hipMemcpy2DAsync(ctx->hs.khouot, ctx->hs.P * sizeof(int32_t), src, W * sizeof(int32_t),
                                 W * sizeof(int32_t), H, hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

static double *boundary_ptr(hd_ctx *ctx, int which, size_t *n) {
    const size_t M = ctx->hs.M, N = ctx->hs.N;
    switch (which) {
        case HD_B_UBT: *n = M + 3; return ctx->hs.ubt;
        case HD_B_UBP: *n = M + 3; return ctx->hs.ubp;
        case HD_B_VBT: *n = N + 3; return ctx->hs.vbt;
        case HD_B_VBD: *n = N + 3; return ctx->hs.vbd;
        case HD_B_HI: *n = 2 * (M + N + 6); return ctx->hs.hi;
        default: *n = 0; return nullptr;
    }
}

int hd_copy_out_boundary(hd_ctx *ctx, int which, double *dst, size_t n) {
    if (!ctx || !dst) return set_err(ctx, HD_E_INVALID, "NULL argument");
    size_t want;
    double *src = boundary_ptr(ctx, which, &want);
    if (!src || n != want) return set_err(ctx, HD_E_INVALID, "bad boundary id or size");
    HD_HIP(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_copy_in_boundary(hd_ctx *ctx, int which, const double *src, size_t n) {
    if (!ctx || !src) return set_err(ctx, HD_E_INVALID, "NULL argument");
    size_t want;
    double *dst = boundary_ptr(ctx, which, &want);
    if (!dst || n != want) return set_err(ctx, HD_E_INVALID, "bad boundary id or size");
    HD_HIP(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_get_hmax(hd_ctx *ctx, double out[4]) {
    if (!ctx || !out) return set_err(ctx, HD_E_INVALID, "NULL argument");
    HD_HIP(ctx, hipMemcpyAsync(out, ctx->hs.hmax, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_set_hmax(hd_ctx *ctx, const double in[4]) {
    if (!ctx || !in) return set_err(ctx, HD_E_INVALID, "NULL argument");
    HD_HIP(ctx, hipMemcpyAsync(ctx->hs.hmax, in, 4 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_get_segments(hd_ctx *ctx, int dir, int32_t *counts, size_t ncounts, int32_t *first, int32_t *last,
                    size_t cap, int64_t *total) {
    if (!ctx || (dir != 0 && dir != 1)) return set_err(ctx, HD_E_INVALID, "bad argument");
    const DevState &s = ctx->hs;
    const size_t nlines = dir == 0 ? (size_t)s.N + 3 : (size_t)s.M + 3;
    const size_t lcap = dir == 0 ? s.capH : s.capV;
    if (counts && ncounts != nlines) return set_err(ctx, HD_E_INVALID, "counts size must be the line count");
    std::vector<int32_t> n(nlines), f(nlines * lcap), l(nlines * lcap);
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HD_HIP(ctx, hipMemcpy(n.data(), dir == 0 ? s.segH_n : s.segV_n, nlines * 4, hipMemcpyDeviceToHost));
    HD_HIP(ctx, hipMemcpy(f.data(), dir == 0 ? s.segH_f : s.segV_f, nlines * lcap * 4, hipMemcpyDeviceToHost));
    HD_HIP(ctx, hipMemcpy(l.data(), dir == 0 ? s.segH_l : s.segV_l, nlines * lcap * 4, hipMemcpyDeviceToHost));
    // lines outside [2, N] / [2, M] carry no segments
    const int lo = 2, hi = dir == 0 ? s.N : s.M;
    int64_t k = 0;
    for (size_t line = 0; line < nlines; line++) {
        const int cnt = ((int)line >= lo && (int)line <= hi) ? n[line] : 0;
        if (counts) counts[line] = cnt;
        for (int q = 0; q < cnt; q++, k++)
            if ((size_t)k < cap && first && last) {
                first[k] = f[line * lcap + q];
                last[k] = l[line * lcap + q];
            }
    }
    if (total) *total = k;
    return HD_OK;
}

int hd_mass_balance(hd_ctx *ctx, double *volume, double *inflow) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    // the wet mask and htaiz come from hd_init
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    LAUNCH(ctx, "k_volume", k_volume<<<1, 1024, 0, ctx->stream>>>(ctx->ds));
    double acc[2];
    HD_HIP(ctx, hipMemcpyAsync(acc, ctx->hs.acc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (inflow) *inflow = acc[0];
    if (volume) *volume = acc[1];
    return HD_OK;
}

int hd_destroy(hd_ctx *ctx) {
    if (!ctx) return HD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    prof_resolve(ctx);
    drop_graphs(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (void *p : ctx->allocs) (void)hipFree(p);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HD_OK;
}

const char *hd_last_error(hd_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int hd_set_profiling(hd_ctx *ctx, int on) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    int rc = prof_resolve(ctx);
    if (rc) return rc;
    ctx->prof = on != 0;
    return HD_OK;
}

int hd_stage_stats(hd_ctx *ctx, char *buf, size_t n, int reset) {
    if (!ctx || !buf || n == 0) return set_err(ctx, HD_E_INVALID, "bad argument");
    int rc = prof_resolve(ctx);
    if (rc) return rc;
    std::string out = "{";
    bool firstk = true;
    for (auto &kv : ctx->stage_stats) {
        char item[256];
        std::snprintf(item, sizeof(item), "%s\"%s\": [%ld, %.6f]", firstk ? "" : ", ", kv.first.c_str(),
                      kv.second.first, kv.second.second);
        out += item;
        firstk = false;
    }
    out += "}";
    if (out.size() + 1 > n) return set_err(ctx, HD_E_INVALID, "buffer too small");
    std::memcpy(buf, out.c_str(), out.size() + 1);
    if (reset) ctx->stage_stats.clear();
    return HD_OK;
}

int hd_device_info(char *buf, size_t n) {
    if (!buf || n == 0) return HD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        std::snprintf(buf, n, "no HIP device");
        return HD_E_NODEV;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return HD_E_HIP;
    std::snprintf(buf, n, "%s %s CUs=%d L2=%d MB HBM=%.1f GB", p.name, p.gcnArchName, p.multiProcessorCount,
                  p.l2CacheSize >> 20, p.totalGlobalMem / 1e9);
    return HD_OK;
}

// ---------------------------------------------------------------------------
// Batches of scenarios and whole-step graphs
// ---------------------------------------------------------------------------
// Switches for batches of more than one scenario and whole-step graphs.
constexpr bool kBatchEnabled = true;
constexpr bool kGraphEnabled = true;

int hd_batch_create(const hd_mesh *meshes, int n, const hd_coeffs *coeffs, const hd_options *ops,
                    const hd_config *cfg, hd_ctx **out) {
    if (n > 1 && !kBatchEnabled) {
        if (out) *out = nullptr;
        return set_err(nullptr, HD_E_UNSUPPORTED, "batches of more than one scenario are switched off in this build");
    }
    return create_ctx(meshes, n, coeffs, ops, cfg, out);
}

int hd_batch_size(hd_ctx *ctx, int *n) {
    if (!ctx || !n) return set_err(ctx, HD_E_INVALID, "NULL argument");
    *n = ctx->nscen;
    return HD_OK;
}

static const DevState *scen_desc(hd_ctx *ctx, int scen) {
    if (scen < 0 || scen >= ctx->nscen) return nullptr;
    return scen == 0 ? &ctx->hs : &ctx->hss[scen - 1];
}

int hd_batch_copy_out(hd_ctx *ctx, int scen, int field, double *dst, size_t n) {
    if (!ctx || !dst) return set_err(ctx, HD_E_INVALID, "NULL argument");
    const DevState *s = scen_desc(ctx, scen);
    if (!s) return set_err(ctx, HD_E_INVALID, "scenario index out of range");
    const size_t W = s->M + 3, H = s->N + 3;
    if (field == HD_F_KHOUOT) {
        if (n != W * H) return set_err(ctx, HD_E_INVALID, "bad size");
        std::vector<int32_t> tmp(n);
        HD_HIP(ctx, hipMemcpy2DAsync(tmp.data(), W * sizeof(int32_t), s->khouot, s->P * sizeof(int32_t),
                                     W * sizeof(int32_t), H, hipMemcpyDeviceToHost, ctx->stream));
        HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t q = 0; q < n; q++) dst[q] = tmp[q];
        return HD_OK;
    }
    double *src = field_ptr(*s, field);
    if (!src || n != W * H) return set_err(ctx, HD_E_INVALID, "bad field id or size");
    HD_HIP(ctx, hipMemcpy2DAsync(dst, W * sizeof(double), src, s->P * sizeof(double), W * sizeof(double), H,
                                 hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_batch_copy_in(hd_ctx *ctx, int scen, int field, const double *src, size_t n) {
    if (!ctx || !src) return set_err(ctx, HD_E_INVALID, "NULL argument");
    const DevState *s = scen_desc(ctx, scen);
    if (!s) return set_err(ctx, HD_E_INVALID, "scenario index out of range");
    double *dst = field_ptr(*s, field);
    const size_t W = s->M + 3, H = s->N + 3;
    if (!dst || n != W * H) return set_err(ctx, HD_E_INVALID, "bad field id or size");
    HD_HIP(ctx, hipMemcpy2DAsync(dst, s->P * sizeof(double), src, W * sizeof(double), W * sizeof(double), H,
                                 hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HD_OK;
}

int hd_footprint(const hd_mesh *mesh, const hd_config *cfg, int nscen, size_t *bytes) {
    if (!mesh || !bytes || nscen < 1 || mesh->M < 4 || mesh->N < 4)
        return set_err(nullptr, HD_E_INVALID, "bad argument");
    *bytes = hd::footprint(mesh, cfg ? cfg->dense_layout : 0, nscen);
    return HD_OK;
}

int hd_memory_bytes(hd_ctx *ctx, size_t *bytes) {
    if (!ctx || !bytes) return set_err(ctx, HD_E_INVALID, "NULL argument");
    *bytes = ctx->bytes;
    return HD_OK;
}

// ---------------------------------------------------------------------------
// Checkpoint / restart
// ---------------------------------------------------------------------------
namespace {

struct SnapHeader {
    char magic[8];  // "HDSNAP3"
    int32_t M, N, P, H, capH, capV, nscen, nbuf;
    double t;
    uint64_t bytes;
    uint64_t sum;  // snap_sum of the header (with sum = 0) and the payload
};

// FNV-1a over 8-byte words (then the tail bytes), continuing from h: a
// flipped or truncated byte of a checkpoint file changes it.
uint64_t fnv_words(const char *p, size_t n, uint64_t h) {
    size_t k = 0;
    for (; k + 8 <= n; k += 8) {
        uint64_t w;
        std::memcpy(&w, p + k, 8);
        h = (h ^ w) * 0x100000001b3ull;
    }
    for (; k < n; k++) h = (h ^ (unsigned char)p[k]) * 0x100000001b3ull;
    return h;
}

uint64_t snap_sum(SnapHeader h, const char *payload, size_t n) {
    h.sum = 0;
    return fnv_words(payload, n, fnv_words(reinterpret_cast<const char *>(&h), sizeof(h), 0xcbf29ce484222325ull));
}

// Every buffer whose contents carry over from one step to the next.
std::vector<std::pair<void *, size_t>> snap_buffers(hd_ctx *ctx) {
    std::vector<std::pair<void *, size_t>> b;
    for (int k = 0; k < ctx->nscen; k++) {
        DevState &s = k == 0 ? ctx->hs : ctx->hss[k - 1];
        const size_t cells = (size_t)s.P * s.H, fb = cells * sizeof(double);
        for (double *f : {s.u, s.v, s.z, s.t_u, s.t_v, s.t_z, s.H_moi, s.Htdu, s.Htdv, s.FS, s.tFS, s.VTH, s.Kx, s.Ky,
                          s.Qbx, s.Qby, s.dH, s.htaiz, s.htaiz_bd, s.Kx1, s.Ky1})
            b.push_back({f, fb});
        b.push_back({s.khouot, cells * sizeof(int)});
        b.push_back({s.segidH, cells * sizeof(int)});
        b.push_back({s.segidV, cells * sizeof(int)});
        b.push_back({s.segflV, 2 * cells * sizeof(int)});
        b.push_back({s.hi, 2 * (size_t)(s.M + s.N + 6) * sizeof(double)});
        b.push_back({s.hmax, 4 * sizeof(double)});
        b.push_back({s.ubt, (s.M + 3) * sizeof(double)});
        b.push_back({s.ubp, (s.M + 3) * sizeof(double)});
        b.push_back({s.vbt, (s.N + 3) * sizeof(double)});
        b.push_back({s.vbd, (s.N + 3) * sizeof(double)});
        b.push_back({s.segH_n, (s.N + 3) * sizeof(int)});
        b.push_back({s.segH_f, (size_t)(s.N + 3) * s.capH * sizeof(int)});
        b.push_back({s.segH_l, (size_t)(s.N + 3) * s.capH * sizeof(int)});
        b.push_back({s.segV_n, (s.M + 3) * sizeof(int)});
        b.push_back({s.segV_f, (size_t)(s.M + 3) * s.capV * sizeof(int)});
        b.push_back({s.segV_l, (size_t)(s.M + 3) * s.capV * sizeof(int)});
        b.push_back({s.acc, 8 * sizeof(double)});
    }
    return b;
}

SnapHeader snap_header(hd_ctx *ctx, double t, size_t nbuf, size_t bytes) {
    SnapHeader h{};
    std::memcpy(h.magic, "HDSNAP3", 8);
    const DevState &s = ctx->hs;
    h.M = s.M; h.N = s.N; h.P = s.P; h.H = s.H; h.capH = s.capH; h.capV = s.capV;
    h.nscen = ctx->nscen;
    h.nbuf = (int32_t)nbuf;
    h.t = t;
    h.bytes = bytes;
    return h;
}

}  // namespace

int hd_snapshot_size(hd_ctx *ctx, size_t *bytes) {
    if (!ctx || !bytes) return set_err(ctx, HD_E_INVALID, "NULL argument");
    size_t n = sizeof(SnapHeader);
    for (auto &b : snap_buffers(ctx)) n += b.second;
    *bytes = n;
    return HD_OK;
}

int hd_snapshot(hd_ctx *ctx, void *buf, size_t n, double t) {
    if (!ctx || !buf) return set_err(ctx, HD_E_INVALID, "NULL argument");
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    size_t need = 0;
    int rc = hd_snapshot_size(ctx, &need);
    if (rc) return rc;
    if (n < need) return set_err(ctx, HD_E_INVALID, "snapshot buffer too small");
    auto bufs = snap_buffers(ctx);
    SnapHeader h = snap_header(ctx, t, bufs.size(), need);
    char *const base = static_cast<char *>(buf);
    char *p = base + sizeof(h);
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &b : bufs) {
        HD_HIP(ctx, hipMemcpy(p, b.first, b.second, hipMemcpyDeviceToHost));
        p += b.second;
    }
    h.sum = snap_sum(h, base + sizeof(h), need - sizeof(h));
    std::memcpy(base, &h, sizeof(h));
    return HD_OK;
}

int hd_restore(hd_ctx *ctx, const void *buf, size_t n, double *t) {
    if (!ctx || !buf) return set_err(ctx, HD_E_INVALID, "NULL argument");
    if (n < sizeof(SnapHeader)) return set_err(ctx, HD_E_INVALID, "snapshot too short");
    SnapHeader h;
    std::memcpy(&h, buf, sizeof(h));
    auto bufs = snap_buffers(ctx);
    size_t need = 0;
    int rc = hd_snapshot_size(ctx, &need);
    if (rc) return rc;
    const SnapHeader want = snap_header(ctx, h.t, bufs.size(), need);
    if (std::memcmp(h.magic, want.magic, 8) || h.M != want.M || h.N != want.N || h.P != want.P || h.H != want.H ||
        h.capH != want.capH || h.capV != want.capV || h.nscen != want.nscen || h.nbuf != want.nbuf ||
        h.bytes != need || n < need)
        return set_err(ctx, HD_E_INVALID, "snapshot does not match this context's mesh and layout");
    const char *p = static_cast<const char *>(buf) + sizeof(h);
    if (snap_sum(h, p, need - sizeof(h)) != h.sum)
        return set_err(ctx, HD_E_INVALID, "snapshot checksum mismatch (damaged or truncated)");
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &b : bufs) {
        HD_HIP(ctx, hipMemcpy(b.first, p, b.second, hipMemcpyHostToDevice));
        p += b.second;
    }
    ctx->inited = true;
    if (t) *t = h.t;
    return HD_OK;
}

int hd_save_checkpoint(hd_ctx *ctx, const char *path, double t) {
    if (!ctx || !path) return set_err(ctx, HD_E_INVALID, "NULL argument");
    size_t n = 0;
    int rc = hd_snapshot_size(ctx, &n);
    if (rc) return rc;
    std::vector<char> buf(n);
    rc = hd_snapshot(ctx, buf.data(), n, t);
    if (rc) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_err(ctx, HD_E_INVALID, std::string("cannot open ") + path);
    const size_t w = std::fwrite(buf.data(), 1, n, f);
    const int c = std::fclose(f);
    if (w != n || c != 0) return set_err(ctx, HD_E_INVALID, std::string("write failed: ") + path);
    return HD_OK;
}

int hd_load_checkpoint(hd_ctx *ctx, const char *path, double *t) {
    if (!ctx || !path) return set_err(ctx, HD_E_INVALID, "NULL argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_err(ctx, HD_E_INVALID, std::string("cannot open ") + path);
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + r);
    std::fclose(f);
    return hd_restore(ctx, buf.data(), buf.size(), t);
}

int hd_autotune(hd_ctx *ctx, char *report, size_t n) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (!ctx->inited) return set_err(ctx, HD_E_STATE, "hd_init has not run");
    std::string r = "{";
    drop_graphs(ctx);
    int rc = hd::tune_line_solvers(ctx, ctx->ds, ctx->hs, ctx->nscen, r);
    if (rc) return rc;
    // choices in the order uz, vz, fs_j, fs_i
    r += r.size() > 1 ? ", \"chosen\": [" : "\"chosen\": [";
    const int order[4][2] = {{1, 1}, {0, 1}, {1, 0}, {0, 0}};
    for (int k = 0; k < 4; k++) {
        const int kind = ctx->solver_kind[order[k][0]][order[k][1]];
        const int param = ctx->solver_param[order[k][0]][order[k][1]];
        char buf[48];
        std::snprintf(buf, sizeof(buf), "%s\"%s/%d\"", k ? ", " : "",
                      kind == HD_SOLVER_CR ? "cr"
                                           : (kind == HD_SOLVER_PARTITION ? "part"
                                                                          : (kind == HD_SOLVER_SEGMENTS ? "seg" : "default")),
                      param);
        r += buf;
    }
    r += "]}";
    if (report && n) std::snprintf(report, n, "%s", r.c_str());
    return HD_OK;
}

int hd_set_solver(hd_ctx *ctx, int dir, int hydraulic, int kind, int param) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (dir != 0 && dir != 1) return set_err(ctx, HD_E_INVALID, "dir must be 0 (VZ) or 1 (UZ)");
    bool ok = false;
    if (kind == HD_SOLVER_DEFAULT) ok = param == 0;
    if (kind == HD_SOLVER_PARTITION) ok = param == 0 || param == 2 || param == 4 || param == 8;
    if (kind == HD_SOLVER_CR) ok = param >= 0 && param <= 512 && param % 64 == 0;
    if (kind == HD_SOLVER_SEGMENTS)
        ok = param == 0 && (ctx->cfg.form == HD_FORM_FP64_PLAIN || ctx->cfg.form == HD_FORM_FP64_FAST);
    if (!ok)
        return set_err(ctx, HD_E_INVALID,
                       "solver: DEFAULT with 0, PARTITION with K in {0, 2, 4, 8}, CR with a workgroup size "
                       "that is a multiple of 64 up to 512 (0 = default), or SEGMENTS with 0 (the FP64 forms)");
    ctx->solver_kind[dir][hydraulic ? 1 : 0] = kind;
    ctx->solver_param[dir][hydraulic ? 1 : 0] = param;
    drop_graphs(ctx);
    return HD_OK;
}

int hd_set_solver_threads(hd_ctx *ctx, int dir, int hydraulic, int threads) {
    return hd_set_solver(ctx, dir, hydraulic, HD_SOLVER_CR, threads);
}

int hd_gtsv_batch(const double *dl, const double *d, const double *du, const double *b, double *x, int m, int nsys,
                  int ld, int precision, void *stream) {
    if (!dl || !d || !du || !b || !x) return set_err(nullptr, HD_E_INVALID, "hd_gtsv_batch: NULL buffer");
    if (m < 1 || m > 4096 || nsys < 0 || (nsys > 1 && ld < m) || (precision != 0 && precision != 1))
        return set_err(nullptr, HD_E_INVALID, "hd_gtsv_batch: need 1 <= m <= 4096, nsys >= 0, ld >= m, precision 0/1");
    if (nsys == 0) return HD_OK;
    hd_ctx c;  // launch settings and error text only; owns nothing
    c.stream = (hipStream_t)stream;
    c.cfg.form = precision ? HD_FORM_FP32_REFINE : HD_FORM_FP64_FAST;
    int rc = hd::solve_batch(&c, dl, d, du, b, x, m, nsys, nsys > 1 ? ld : m);
    if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) {
        c.err = std::string("hd_gtsv_batch: ") + hipGetErrorString(hipGetLastError());
        rc = HD_E_HIP;
    }
    if (rc) return set_err(nullptr, rc, c.err);
    return HD_OK;
}

int hd_set_graph(hd_ctx *ctx, int on) {
    if (!ctx) return set_err(nullptr, HD_E_INVALID, "ctx is NULL");
    if (on && !kGraphEnabled) return set_err(ctx, HD_E_UNSUPPORTED, "whole-step graphs are switched off in this build");
    if ((on != 0) != ctx->use_graph) drop_graphs(ctx);
    ctx->use_graph = on != 0;
    return HD_OK;
}

}  // extern "C"
