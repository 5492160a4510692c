// hd_compat.hip - the reference's host surface (engine.h / loader.h,
// main.cpp's init sequence) over the gfx950 library.  See
// include/hydist_compat.h for what a caller can observe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "hd_internal.h"
#include "hydist_compat.h"

namespace {

struct Entry {
    hd_ctx *ctx;
    Argument_Pointers host;
};

std::mutex g_mu;
std::map<const void *, Entry> g_by_dev;   // device Argument_Pointers* -> context
std::map<const void *, hd_ctx *> g_by_u;  // host struct's u pointer -> context

[[noreturn]] void fatal(const char *what, hd_ctx *ctx) {
    std::cout << "error: " << what << ": " << hd_last_error(ctx) << std::endl;
    std::abort();
}

void check(int rc, const char *what, hd_ctx *ctx) {
    if (rc != HD_OK) fatal(what, ctx);
}

// HYDIST_FORM: plain | fast | fp32, or the digits 0-2 (HD_FORM_*); unset
// means plain.  Anything else ends the program with status 2 before any
// device work, so a misspelt form never runs silently as another.
int form_from_env(const char *v) {
    if (!v || !*v) return HD_FORM_FP64_PLAIN;
    const std::string s(v);
    if (s == "plain" || s == "0") return HD_FORM_FP64_PLAIN;
    if (s == "fast" || s == "1") return HD_FORM_FP64_FAST;
    if (s == "fp32" || s == "2") return HD_FORM_FP32_REFINE;
    std::cout << "error: HYDIST_FORM=" << s << ": expected plain, fast or fp32 (or 0, 1, 2)" << std::endl;
    std::exit(2);
}

// The drop-in path keeps the reference's literal end test of the
// open-start branch (dkBienQ_1, UVZSolver_multithread.cpp:438) unless
// HYDIST_Q11=defined asks for the end side's own flag.
int quirks_from_env(const char *v) {
    if (!v || !*v || std::string(v) == "literal") return HD_QUIRK_Q11_LITERAL;
    if (std::string(v) == "defined") return 0;
    std::cout << "error: HYDIST_Q11=" << v << ": expected literal or defined" << std::endl;
    std::exit(2);
}

hd_ctx *lookup(const void *key) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_by_dev.find(key);
    if (it != g_by_dev.end()) return it->second.ctx;
    return nullptr;
}

// Copies a Constant_Coeffs that may live in device memory (the reference
// keeps it there, main.cpp:231-232).
Constant_Coeffs read_coeffs(const Constant_Coeffs *p) {
    Constant_Coeffs c{};
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) == hipSuccess && attr.type == hipMemoryTypeDevice) {
        if (hipMemcpy(&c, p, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) fatal("read coeffs", nullptr);
    } else {
        (void)hipGetLastError();
        c = *p;
    }
    return c;
}

void set_coeffs(hd_ctx *ctx, const Constant_Coeffs &c) {
    ctx->c = c;
    ctx->hs.c = c;
    check(hd::sync_desc(ctx), "upload coeffs", ctx);
}

Argument_Pointers make_args(hd_ctx *ctx) {
    const hd::DevState &s = ctx->hs;
    Argument_Pointers a{};
    a.M = s.M;
    a.N = s.N;
    a.khouot = s.khouot;
    a.daui = s.segH_f;
    a.cuoii = s.segH_l;
    a.moci = s.segH_n;
    a.dauj = s.segV_f;
    a.cuoij = s.segV_l;
    a.mocj = s.segV_n;
    a.h = const_cast<double *>(s.h);
    a.u = s.u; a.v = s.v; a.z = s.z;
    a.t_u = s.t_u; a.t_v = s.t_v; a.t_z = s.t_z;
    a.Htdu = s.Htdu; a.Htdv = s.Htdv; a.H_moi = s.H_moi;
    a.htaiz = s.htaiz; a.htaiz_bd = s.htaiz_bd;
    a.ubt = s.ubt; a.ubp = s.ubp; a.vbt = s.vbt; a.vbd = s.vbd;
    a.hsnham = const_cast<double *>(s.hsnham);
    a.VISCOIDX = const_cast<double *>(s.visc);
    a.Kx1 = s.Kx1; a.Ky1 = s.Ky1;
    a.Tsxw = const_cast<double *>(s.Tsxw);
    a.Tsyw = const_cast<double *>(s.Tsyw);
    a.bc_up = const_cast<double *>(s.bc[HD_SIDE_UP]);
    a.bc_down = const_cast<double *>(s.bc[HD_SIDE_DOWN]);
    a.bc_left = const_cast<double *>(s.bc[HD_SIDE_LEFT]);
    a.bc_right = const_cast<double *>(s.bc[HD_SIDE_RIGHT]);
    a.hi = s.hi;
    a.FS = s.FS; a.tFS = s.tFS;
    a.CC_u = const_cast<double *>(s.CC[HD_SIDE_UP]);
    a.CC_d = const_cast<double *>(s.CC[HD_SIDE_DOWN]);
    a.CC_l = const_cast<double *>(s.CC[HD_SIDE_LEFT]);
    a.CC_r = const_cast<double *>(s.CC[HD_SIDE_RIGHT]);
    a.VTH = s.VTH; a.Kx = s.Kx; a.Ky = s.Ky;
    a.Fw = const_cast<double *>(s.Fw);
    a.Qbx = s.Qbx; a.Qby = s.Qby; a.dH = s.dH;
    return a;
}

// loader.cpp:378-389: interior rows 1..height, columns 1..width.
template <typename T>
void save_grid(const std::vector<T> &a, int width, int height, const std::string &name, double scale) {
    std::ofstream ofs(name);
    if (!ofs) {
        std::cout << "cannot open file " << name << std::endl;
        return;
    }
    for (int i = 1; i <= height; i++) {
        for (int j = 1; j <= width; j++) ofs << a[(size_t)i * (width + 3) + j] * scale << " ";
        ofs << std::endl;
    }
}

}  // namespace

void Load_coeffs(Constant_Coeffs &ce) { hd_load_coeffs(&ce); }

template <typename T>
std::pair<int, int> load_file(const char *filename, std::vector<T> &arr) {
    std::ifstream ifs(filename);
    if (!ifs) std::cout << "cannot open" << filename << std::endl;
    std::string line;
    int M = 0, N = 0;
    while (std::getline(ifs, line)) {
        N++;
        std::stringstream ls(line);
        T x;
        M = 0;
        while (ls >> x) {
            M++;
            arr.push_back(x);
        }
    }
    return std::pair<int, int>(N, M);
}

template std::pair<int, int> load_file<double>(const char *, std::vector<double> &);
template std::pair<int, int> load_file<float>(const char *, std::vector<float> &);
template std::pair<int, int> load_file<int>(const char *, std::vector<int> &);

std::pair<int, int> load_inputs(std::string dir, std::vector<DOUBLE> &h, std::vector<DOUBLE> &hsnham,
                                std::vector<DOUBLE> &VISCOIDX, std::vector<DOUBLE> &Fw,
                                std::vector<DOUBLE> &bc_up, std::vector<DOUBLE> &bc_right,
                                std::vector<DOUBLE> &bc_left, std::vector<DOUBLE> &bc_down,
                                std::vector<DOUBLE> &CC_u, std::vector<DOUBLE> &CC_d, std::vector<DOUBLE> &CC_l,
                                std::vector<DOUBLE> &CC_r, std::vector<int> &bienQ, std::vector<int> &boundary_type,
                                int *total_time) {
    std::pair<int, int> size = load_file<DOUBLE>((dir + "bandodosau.txt").c_str(), h);
    load_file<DOUBLE>((dir + "hsnham.txt").c_str(), hsnham);
    load_file<DOUBLE>((dir + "hsnhotroiA.txt").c_str(), VISCOIDX);
    load_file<DOUBLE>((dir + "Fw_map.txt").c_str(), Fw);
    load_file<DOUBLE>((dir + "bientren.txt").c_str(), bc_up);
    std::pair<int, int> shaped = load_file<DOUBLE>((dir + "bienduoi.txt").c_str(), bc_down);
    load_file<DOUBLE>((dir + "bientrai.txt").c_str(), bc_left);
    load_file<DOUBLE>((dir + "bienphai.txt").c_str(), bc_right);
    load_file<DOUBLE>((dir + "FS_bientren.txt").c_str(), CC_u);
    load_file<DOUBLE>((dir + "FS_bienduoi.txt").c_str(), CC_d);
    load_file<DOUBLE>((dir + "FS_bientrai.txt").c_str(), CC_l);
    load_file<DOUBLE>((dir + "FS_bienphai.txt").c_str(), CC_r);
    load_file<int>((dir + "boundary_type.txt").c_str(), boundary_type);
    for (int v : boundary_type) bienQ.push_back(v == 2 ? 1 : 0);
    *total_time = shaped.first;  // rows of bienduoi.txt (loader.cpp:165)
    return size;
}

Argument_Pointers attribute_arrays_memory_alloc(int device, Host_arrays &ap, Argument_Pointers **device_arg_ptr) {
    const int M = ap.M, N = ap.N;
    const size_t cells = (size_t)(M + 3) * (N + 3);
    if (ap.h.size() < cells || ap.hsnham.size() < cells || ap.VISCOIDX.size() < cells) {
        std::cout << "error: input grids smaller than (N+3)*(M+3)" << std::endl;
        std::abort();
    }
    hd_mesh mesh{};
    mesh.M = M;
    mesh.N = N;
    mesh.h = ap.h.data();
    mesh.hsnham = ap.hsnham.data();
    mesh.VISCOIDX = ap.VISCOIDX.data();
    mesh.Fw = ap.Fw.size() >= cells ? ap.Fw.data() : nullptr;
    // initial condition only when all of u, v, z, khouot are given (loader.cpp:261-266)
    if (ap.u.size() >= cells && ap.v.size() >= cells && ap.z.size() >= cells && !ap.khouot.empty()) {
        mesh.u0 = ap.u.data();
        mesh.v0 = ap.v.data();
        mesh.z0 = ap.z.data();
    }
    // FS is re-allocated after the optional copy in the reference (loader.cpp:275-300): starts at 0
    const int T = ap.total_time > 1 ? ap.total_time : 2;
    mesh.total_time = T;
    std::vector<DOUBLE> *series[4] = {&ap.bc_up, &ap.bc_down, &ap.bc_left, &ap.bc_right};
    std::vector<DOUBLE> *cc[4] = {&ap.CC_u, &ap.CC_d, &ap.CC_l, &ap.CC_r};
    static thread_local std::vector<DOUBLE> padded[8];
    for (int s = 0; s < 4; s++) {
        mesh.boundary_type[s] = s < (int)ap.boundary_type.size() ? ap.boundary_type[s] : 0;
        for (int k = 0; k < 2; k++) {
            std::vector<DOUBLE> &src = k == 0 ? *series[s] : *cc[s];
            std::vector<DOUBLE> &dst = padded[2 * s + k];
            dst.assign(T, 0.0);
            for (int q = 0; q < T; q++)
                dst[q] = src.empty() ? 0.0 : src[std::min<size_t>(q, src.size() - 1)];
        }
        mesh.bc[s] = padded[2 * s].data();
        mesh.CC[s] = padded[2 * s + 1].data();
    }
    hd_coeffs c;
    hd_load_coeffs(&c);
    hd_options ops{};
    ops.M = M;
    ops.N = N;
    ops.ketdinh = true;
    ops.total_time = T;
    ops.sediment_start = 1e300;
    ops.bed_change_start = 1e300;
    hd_config cfg{};
    cfg.form = form_from_env(std::getenv("HYDIST_FORM"));
    cfg.quirks = quirks_from_env(std::getenv("HYDIST_Q11"));
    cfg.sediment = 1;
    cfg.bed_change = 0;  // the reference never launches BedLoad (engine.cpp:290-293)
    cfg.device = device;
    cfg.dense_layout = 1;
    hd_ctx *ctx = nullptr;
    int rc = hd_create(&mesh, &c, &ops, &cfg, &ctx);
    if (rc) fatal("attribute_arrays_memory_alloc", nullptr);
    Argument_Pointers a = make_args(ctx);
    Argument_Pointers *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(Argument_Pointers)) != hipSuccess ||
        hipMemcpy(d, &a, sizeof(a), hipMemcpyHostToDevice) != hipSuccess)
        fatal("device copy of Argument_Pointers", ctx);
    ctx->allocs.push_back(d);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_by_dev[d] = Entry{ctx, a};
        g_by_u[a.u] = ctx;
    }
    *device_arg_ptr = d;
    return a;
}

Array_Pointers supporting_arrays_alloc(int M, int N, Array_Pointers **device_arr_ptr) {
    // The most recent context with this shape owns the sweep scratch.
    hd_ctx *ctx = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto &kv : g_by_dev)
            if (kv.second.ctx->hs.M == M && kv.second.ctx->hs.N == N) ctx = kv.second.ctx;
    }
    Array_Pointers a{};
    if (ctx) {
        const hd::SweepBufs &B = ctx->sweep;
        a.AA = B.A; a.BB = B.B; a.CC = B.C; a.DD = B.D; a.x = B.X; a.Ap = B.Ap; a.Bp = B.Bp;
        a.SN = B.sysn;
    }
    Array_Pointers *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(Array_Pointers)) != hipSuccess ||
        hipMemcpy(d, &a, sizeof(a), hipMemcpyHostToDevice) != hipSuccess)
        fatal("device copy of Array_Pointers", ctx);
    if (ctx) ctx->allocs.push_back(d);
    *device_arr_ptr = d;
    return a;
}

hd_ctx *hydist_compat_context(const Argument_Pointers *arg) {
    hd_ctx *c = lookup(arg);
    if (c) return c;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_by_u.find(arg->u);
    return it == g_by_u.end() ? nullptr : it->second;
}

void hydist_init(Argument_Pointers *d_arg_ptr, Constant_Coeffs *coeffs) {
    hd_ctx *ctx = lookup(d_arg_ptr);
    if (!ctx) fatal("hydist_init: unknown Argument_Pointers", nullptr);
    set_coeffs(ctx, read_coeffs(coeffs));
    check(hd_init(ctx), "hydist_init", ctx);
    check(hd_sync(ctx), "hydist_init", ctx);
}

void update_boundary_at_t(int M, int N, float t, bool channel, int total_time, Argument_Pointers *d_arg_ptr,
                          Constant_Coeffs *coeffs) {
    (void)M;
    (void)N;
    (void)total_time;
    (void)coeffs;
    if (channel) return;  // engine.cpp:24-30: no boundary in narrow-channel mode
    hd_ctx *ctx = lookup(d_arg_ptr);
    if (!ctx) fatal("update_boundary_at_t: unknown Argument_Pointers", nullptr);
    check(hd_update_boundary(ctx, t), "update_boundary_at_t", ctx);
}

void synch_and_check() {
    std::vector<hd_ctx *> all;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto &kv : g_by_dev) all.push_back(kv.second.ctx);
    }
    for (hd_ctx *c : all) check(hd_sync(c), "synch_and_check", c);
}

void save_result(Argument_Pointers h, int t) {
    hd_ctx *ctx = hydist_compat_context(&h);
    if (!ctx) fatal("save_result: unknown context", nullptr);
    const size_t n = (size_t)(h.M + 3) * (h.N + 3);
    std::vector<double> u(n), v(n), z(n);
    check(hd_copy_out(ctx, HD_F_U, u.data(), n), "save_result", ctx);
    check(hd_copy_out(ctx, HD_F_V, v.data(), n), "save_result", ctx);
    check(hd_copy_out(ctx, HD_F_Z, z.data(), n), "save_result", ctx);
    const std::string ts = std::to_string(t);
    save_grid(u, h.M, h.N, "Outputs/U/u_" + ts + ".txt", 1.0);
    save_grid(v, h.M, h.N, "Outputs/V/v_" + ts + ".txt", 1.0);
    save_grid(z, h.M, h.N, "Outputs/Z/z_" + ts + ".txt", 1.0);
}

void save_FS(Argument_Pointers h, int t, DOUBLE ros) {
    hd_ctx *ctx = hydist_compat_context(&h);
    if (!ctx) fatal("save_FS: unknown context", nullptr);
    const size_t n = (size_t)(h.M + 3) * (h.N + 3);
    std::vector<double> fs(n);
    check(hd_copy_out(ctx, HD_F_FS, fs.data(), n), "save_FS", ctx);
    save_grid(fs, h.M, h.N, "Outputs/FS/fs_" + std::to_string(t) + ".txt", ros);
}

void Hydraulic_Calculation(DOUBLE dT, DOUBLE NANGDAY, Argument_Pointers *d_arg_ptr, Array_Pointers *d_arr_ptr,
                           Constant_Coeffs *coeffs, Argument_Pointers h_arg_ptr, Options ops) {
    (void)d_arr_ptr;
    hd_ctx *ctx = lookup(d_arg_ptr);
    if (!ctx) fatal("Hydraulic_Calculation: unknown Argument_Pointers", nullptr);
    Constant_Coeffs c = read_coeffs(coeffs);
    c.dT = dT;
    c.NANGDAY = NANGDAY;
    set_coeffs(ctx, c);
    hd_options &o = ctx->ops;
    o.Tmax = ops.Tmax;
    o.t_start = ops.t_start;
    o.sediment_start = ops.sediment_start;
    o.bed_change_start = ops.bed_change_start;
    o.interval = ops.interval;
    o.kenhhepd = ops.kenhhepd;
    o.kenhhepng = ops.kenhhepng;
    o.ketdinh = ops.ketdinh;
    o.channel = ops.channel;
    o.debug = ops.debug;
    o.plot = ops.plot;
    ctx->hs.ketdinh = ops.ketdinh ? 1 : 0;
    check(hd::sync_desc(ctx), "Hydraulic_Calculation", ctx);
    if (!ctx->inited) check(hd_init(ctx), "Hydraulic_Calculation: init", ctx);
    DOUBLE t = ops.t_start;
    const int Tmax = (int)ops.Tmax;  // engine.cpp:62
    const int interval = ops.interval > 0 ? ops.interval : 1;
    std::cout << "t = " << t << " sediment start = " << ops.sediment_start << std::endl;
    while (t < Tmax) {
        t += 0.5 * dT;
        if (ops.debug) std::cout << t << std::endl;
        check(hd_halfstep_uz(ctx, t), "UZ half-step", ctx);
        if (((int)t % interval == 0) && (t - (int)t == 0)) {
            check(hd_sync(ctx), "sync", ctx);
            save_result(h_arg_ptr, (int)t);
        }
        t += dT * 0.5;
        if (ops.debug) std::cout << t << std::endl;
        check(hd::run_halfstep_vz_a(ctx, ctx->ds, ctx->hs, 1, t), "VZ half-step", ctx);
        if (((int)t % interval == 0) && (t - (int)t == 0)) {
            check(hd_sync(ctx), "sync", ctx);
            save_result(h_arg_ptr, (int)t);
        }
        check(hd::run_halfstep_vz_b(ctx, ctx->ds, ctx->hs, 1, t), "VZ half-step", ctx);
        if (ctx->cfg.sediment && t >= ops.sediment_start && ((int)t % 360 == 0) && (t - (int)t == 0)) {
            check(hd_sync(ctx), "sync", ctx);
            save_FS(h_arg_ptr, (int)t, 2000);  // engine.cpp:283-286
        }
    }
    check(hd_sync(ctx), "Hydraulic_Calculation", ctx);
}
