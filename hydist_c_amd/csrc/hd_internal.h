// hd_internal.h - host-side context and stage sequences shared by the C ABI
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "hd_kernels.h"
#include "hd_sweep.h"

struct hd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hd_coeffs c{};
    hd_options ops{};
    hd_config cfg{};
    hd::DevState hs{};             // descriptor of scenario 0 (host copy)
    std::vector<hd::DevState> hss; // descriptors of scenarios 1..n-1 of a batch
    int nscen = 1;
    hd::DevState *ds = nullptr;    // descriptors of all scenarios (device copy)
    double *clock = nullptr;       // device time of the current half-step
    // line solver of the fast forms per sweep [dir][hydraulic]: the kind
    // (HD_SOLVER_*, HD_SOLVER_DEFAULT = the default rule) and its parameter
    // (unknowns per thread for the partition solver, workgroup size for
    // cyclic reduction; 0 = that solver's default); set by hd_autotune or
    // hd_set_solver
    int solver_kind[2][2] = {{0, 0}, {0, 0}};
    int solver_param[2][2] = {{0, 0}, {0, 0}};
    bool use_graph = false;
    bool capturing = false;
    // whole-step graphs, keyed by which optional stages the step runs
    std::map<int, hipGraphExec_t> graphs;
    hd::SweepBufs sweep{};         // line-sweep buffers
    std::vector<void *> allocs;
    size_t bytes = 0;
    bool inited = false;
    int fault_stage = 0;           // test hook: 1 = fail UZ half-step, 2 = fail VZ
    // NaN/Inf watchdog: a check after every wd_every-th step into wd_flag
    int wd_every = 0;
    long wd_steps = 0;
    bool wd_tripped = false;
    int *wd_flag = nullptr;
    std::string err;
    // stage profiling: events recorded on `stream` around each launch
    bool prof = false;
    std::vector<hipEvent_t> event_pool;
    struct Pending {
        const char *name;
        hipEvent_t e0, e1;
    };
    std::vector<Pending> pending;
    std::map<std::string, std::pair<long, double>> stage_stats;  // name -> (launches, total ms)
};

namespace hd {

struct Launch2D {
    dim3 grid, block;
};

int ctx_alloc(hd_ctx *ctx, void **p, size_t bytes);
double *field_ptr(hd_ctx *ctx, int f);
double *field_ptr(const DevState &s, int f);
Launch2D grid2d(const DevState &s, int nscen);
int sync_desc(hd_ctx *ctx);
int check_launch(hd_ctx *ctx, const char *what);
// Records msg as the library's last error (hd_last_error(NULL)) and
// returns code.
int set_library_error(int code, const std::string &msg);
void drop_graphs(hd_ctx *ctx);
hipEvent_t prof_begin(hd_ctx *ctx);
void prof_end(hd_ctx *ctx, const char *what, hipEvent_t e0);
int prof_resolve(hd_ctx *ctx);

// Launches a kernel (the variadic part), checks the launch and, when
// profiling, brackets it with events on the context stream.
#define LAUNCH(ctx, what, ...)                          \
    do {                                                \
        hipEvent_t e0_ = hd::prof_begin(ctx);           \
        __VA_ARGS__;                                    \
        int rc_ = hd::check_launch((ctx), (what));      \
        if (rc_) return rc_;                            \
        hd::prof_end((ctx), (what), e0_);               \
    } while (0)
// As LAUNCH, storing the status in `rc` instead of returning it.
#define LAUNCH_RC(ctx, rc, what, ...)                   \
    do {                                                \
        hipEvent_t e0_ = hd::prof_begin(ctx);           \
        __VA_ARGS__;                                    \
        (rc) = hd::check_launch((ctx), (what));         \
        if (!(rc)) hd::prof_end((ctx), (what), e0_);    \
    } while (0)
int alloc_scenario(hd_ctx *ctx, const hd_mesh *mesh, DevState &s);
int alloc_shared(hd_ctx *ctx, int nscen);
size_t footprint(const hd_mesh *mesh, int dense, int nscen);
// HD_OK, or HD_E_INVALID (message set) for an unknown form or quirk bit, or
// a grid spacing or time step that is not a positive finite number.
int check_setup(const hd_coeffs *coeffs, const hd_config *cfg);

int run_segmentation(hd_ctx *ctx, const DevState *S, const DevState &s, int n);
int run_init(hd_ctx *ctx, const DevState *S, const DevState &s, int n);
int run_halfstep_uz(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t);
int run_halfstep_vz(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t);
int run_halfstep_vz_a(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t);
int run_halfstep_vz_b(hd_ctx *ctx, const DevState *S, const DevState &s, int n, double t);
int run_sweep(hd_ctx *ctx, const DevState *S, const DevState &s, int n, int dir, int start, int end);

}  // namespace hd
