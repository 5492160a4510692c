// hd_gtsv.hip - one tridiagonal system of any length on the device
// (gtsv_spike_partial_diag_pivot_v1, gtsv.cpp:88-193, splits one system of
// m rows over s workgroups with a level-2 interface system solved in one
// block; same argument meaning here).
//
// The system is cut into chunks of GT_CHUNK rows.  In chunk g the last row
// e_g holds the interface unknown L_g; the rows before it ("interior") form
// a tridiagonal system T_g coupled to L_{g-1} through its first row and to
// L_g through its last.  Solving T_g for three right-hand sides,
//     T_g y = d,   T_g v = -a_first e_1,   T_g w = -c_last e_last,
// every interior unknown is x_k = y_k + v_k L_{g-1} + w_k L_g.  Substituted
// into the rows e_g this leaves one tridiagonal row per chunk in the L (a
// Schur complement, diagonally dominant when the system is), which is
// solved, and the interior follows.  All three interior solves and the
// interface solve are the batched partition solver (solve_batch), reading
// the caller's coefficient arrays in place (chunk g's interior starts at
// row g * GT_CHUNK, so the chunks are systems GT_CHUNK elements apart).
//
// Device memory beyond the caller's arrays ("scratch"): the two spike
// right-hand sides and solutions (4 m doubles) and the interface system
// (5 values per chunk).  hd_gtsv_long allocates it on the caller's stream
// for each call; hd_gtsv_long_noalloc takes it from a fixed device array
// sized for HD_GTSV_NOALLOC_MAX_M rows.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "hd_internal.h"

namespace hd {

constexpr int GT_CHUNK = 4096;
constexpr int GT_MAX_CHUNKS = 4096;  // interface system solved by one workgroup

namespace {

__host__ __device__ inline int chunks_of(int m) { return (m + GT_CHUNK - 1) / GT_CHUNK; }

size_t scratch_doubles(int m) { return 4 * (size_t)m + 5 * (size_t)chunks_of(m); }

// Last row of chunk g.
HD_DEV int chunk_end(int g, int m) { return min((g + 1) * GT_CHUNK, m) - 1; }

// Spike right-hand sides: -a at each chunk's first interior row (chunks
// after the first) and -c at each chunk's last interior row; zero
// elsewhere (interface rows included, which no interior solve reads).
__global__ void k_spike_rhs(const double *__restrict__ dl, const double *__restrict__ du, int m,
                            double *__restrict__ rv, double *__restrict__ rw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int g = k / GT_CHUNK, s = g * GT_CHUNK, e = chunk_end(g, m);
    rv[k] = (k == s && g > 0 && k < e) ? -dl[k] : 0.0;
    rw[k] = (k == e - 1 && k >= s) ? -du[k] : 0.0;
}

// Interface system: row g in L_{g-1}, L_g, L_{g+1}.
__global__ void k_spike_reduce(const double *__restrict__ dl, const double *__restrict__ d,
                               const double *__restrict__ du, const double *__restrict__ b,
                               const double *__restrict__ y, const double *__restrict__ v,
                               const double *__restrict__ w, int m, double *__restrict__ ra,
                               double *__restrict__ rb, double *__restrict__ rc, double *__restrict__ rd) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = chunks_of(m);
    if (g >= P) return;
    const int s = g * GT_CHUNK, e = chunk_end(g, m);
    // x_{e-1} = ym + vm L_{g-1} + wm L_g (the interior's last row, or
    // L_{g-1} itself when the chunk has no interior)
    const bool in = e > s;
    const double ym = in ? y[e - 1] : 0.0, vm = in ? v[e - 1] : 1.0, wm = in ? w[e - 1] : 0.0;
    // row e couples to x_{e-1} through dl[e] (e >= 4095: never the system's
    // first row, whose dl is ignored) and, except in the last chunk, to
    // x_{e+1} through du[e]
    const double a = dl[e], c = g + 1 < P ? du[e] : 0.0;
    double yp = 0.0, vp = 0.0, wp = 0.0;  // x_{e+1} = yp + vp L_g + wp L_{g+1}
    if (g + 1 < P) {
        const int s1 = e + 1, e1 = chunk_end(g + 1, m);
        if (e1 > s1) {
            yp = y[s1];
            vp = v[s1];
            wp = w[s1];
        } else {
            wp = 1.0;
        }
    }
    ra[g] = a * vm;
    rb[g] = d[e] + a * wm + c * vp;
    rc[g] = c * wp;
    rd[g] = b[e] - a * ym - c * yp;
}

// x = y + v L_{g-1} + w L_g on the interior rows (y is already in x), and
// the interface unknowns.
__global__ void k_spike_finish(const double *__restrict__ v, const double *__restrict__ w,
                               const double *__restrict__ L, int m, double *__restrict__ x) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int g = k / GT_CHUNK, e = chunk_end(g, m);
    if (k == e) {
        x[k] = L[g];
    } else {
        const double lm = g > 0 ? L[g - 1] : 0.0;
        x[k] = x[k] + v[k] * lm + w[k] * L[g];
    }
}

// The interior systems of every chunk for one right-hand side rhs into
// out: the full chunks as one batch, the last chunk apart when shorter.
int solve_interiors(hd_ctx *ctx, const double *dl, const double *d, const double *du, const double *rhs, double *out,
                    int m) {
    const int full = m / GT_CHUNK, rem = m - full * GT_CHUNK;
    int rc = HD_OK;
    if (full > 0) rc = solve_batch(ctx, dl, d, du, rhs, out, GT_CHUNK - 1, full, GT_CHUNK);
    if (!rc && rem >= 2) {
        const size_t o = (size_t)full * GT_CHUNK;
        rc = solve_batch(ctx, dl + o, d + o, du + o, rhs + o, out + o, rem - 1, 1, rem - 1);
    }
    return rc;
}

}  // namespace

// The whole solve on ctx->stream with `scratch` (scratch_doubles(m)
// doubles of device memory).
int gtsv_long(hd_ctx *ctx, const double *dl, const double *d, const double *du, const double *b, double *x, int m,
              double *scratch) {
    if (m <= GT_CHUNK) return solve_batch(ctx, dl, d, du, b, x, m, 1, m);
    const int P = chunks_of(m);
    double *rv = scratch, *rw = rv + m, *v = rw + m, *w = v + m;
    double *ra = w + m, *rb = ra + P, *rcc = rb + P, *rd = rcc + P, *L = rd + P;
    hipStream_t st = ctx->stream;
    const int blocks = (m + 255) / 256;
    LAUNCH(ctx, "k_spike_rhs", k_spike_rhs<<<blocks, 256, 0, st>>>(dl, du, m, rv, rw));
    int rc = solve_interiors(ctx, dl, d, du, b, x, m);
    if (!rc) rc = solve_interiors(ctx, dl, d, du, rv, v, m);
    if (!rc) rc = solve_interiors(ctx, dl, d, du, rw, w, m);
    if (rc) return rc;
    LAUNCH(ctx, "k_spike_reduce",
           k_spike_reduce<<<(P + 255) / 256, 256, 0, st>>>(dl, d, du, b, x, v, w, m, ra, rb, rcc, rd));
    rc = solve_batch(ctx, ra, rb, rcc, rd, L, P, 1, P);
    if (rc) return rc;
    LAUNCH(ctx, "k_spike_finish", k_spike_finish<<<blocks, 256, 0, st>>>(v, w, L, m, x));
    return HD_OK;
}

}  // namespace hd

namespace {

// Scratch of hd_gtsv_long_noalloc: one copy per device, held until the
// process ends.  Host code never uses the array's name as an address: each
// device's copy is looked up once (hipGetSymbolAddress, with that device
// current) and the pointer kept per device.
constexpr size_t kNoAllocDoubles = 4 * (size_t)HD_GTSV_NOALLOC_MAX_M +
                                   5 * (size_t)((HD_GTSV_NOALLOC_MAX_M + hd::GT_CHUNK - 1) / hd::GT_CHUNK);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`, looked up on first use.
int device_scratch(double **out, std::string &err) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return HD_E_HIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return HD_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return HD_OK;
}

int gtsv_args(const double *dl, const double *d, const double *du, const double *b, double *x, int m, int precision,
              const char *who, std::string &err) {
    if (!dl || !d || !du || !b || !x) {
        err = std::string(who) + ": NULL buffer";
        return HD_E_INVALID;
    }
    if (m < 1 || m > hd::GT_CHUNK * hd::GT_MAX_CHUNKS || (precision != 0 && precision != 1)) {
        err = std::string(who) + ": need 1 <= m <= " + std::to_string(hd::GT_CHUNK * hd::GT_MAX_CHUNKS) +
              " and precision 0 or 1";
        return HD_E_INVALID;
    }
    return HD_OK;
}

}  // namespace

extern "C" {

int hd_gtsv_long(const double *dl, const double *d, const double *du, const double *b, double *x, int m,
                 int precision, void *stream) {
    hd_ctx c;  // launch settings and error text only; owns nothing
    int rc = gtsv_args(dl, d, du, b, x, m, precision, "hd_gtsv_long", c.err);
    if (rc) return hd::set_library_error(rc, c.err.c_str());
    c.stream = (hipStream_t)stream;
    c.cfg.form = precision ? HD_FORM_FP32_REFINE : HD_FORM_FP64_FAST;
    void *p = nullptr;
    if (m > hd::GT_CHUNK) {
        hipError_t e = hipMallocAsync(&p, hd::scratch_doubles(m) * sizeof(double), c.stream);
        if (e != hipSuccess) return hd::set_library_error(HD_E_NOMEM, (std::string("hd_gtsv_long: hipMallocAsync: ") +
                                                                      hipGetErrorString(e)).c_str());
    }
    rc = hd::gtsv_long(&c, dl, d, du, b, x, m, static_cast<double *>(p));
    if (p) (void)hipFreeAsync(p, c.stream);
    if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) {
        c.err = std::string("hd_gtsv_long: ") + hipGetErrorString(hipGetLastError());
        rc = HD_E_HIP;
    }
    if (rc) return hd::set_library_error(rc, c.err.c_str());
    return HD_OK;
}

int hd_gtsv_long_noalloc(const double *dl, const double *d, const double *du, const double *b, double *x, int m,
                         int precision, void *stream) {
    hd_ctx c;  // launch settings and error text only; owns nothing
    int rc = gtsv_args(dl, d, du, b, x, m, precision, "hd_gtsv_long_noalloc", c.err);
    if (!rc && m > HD_GTSV_NOALLOC_MAX_M) {
        c.err = "hd_gtsv_long_noalloc: m = " + std::to_string(m) + " exceeds the fixed scratch's " +
                std::to_string(HD_GTSV_NOALLOC_MAX_M) + " rows (use hd_gtsv_long)";
        rc = HD_E_INVALID;
    }
    if (rc) return hd::set_library_error(rc, c.err.c_str());
    double *p = nullptr;
    rc = device_scratch(&p, c.err);
    if (rc) return hd::set_library_error(rc, c.err.c_str());
    c.stream = (hipStream_t)stream;
    c.cfg.form = precision ? HD_FORM_FP32_REFINE : HD_FORM_FP64_FAST;
    rc = hd::gtsv_long(&c, dl, d, du, b, x, m, p);
    if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) {
        c.err = std::string("hd_gtsv_long_noalloc: ") + hipGetErrorString(hipGetLastError());
        rc = HD_E_HIP;
    }
    if (rc) return hd::set_library_error(rc, c.err.c_str());
    return HD_OK;
}

}  // extern "C"
