// hd_server.hip - serving loop for what-if runs on one GPU.
//
// The reference runs one scenario per process, start to finish
// (main.cpp:138-309).  The server instead takes runs as they arrive
// (different meshes, hydrographs, roughness, durations, forms), keeps each
// admitted run's state resident in HBM in its own context, and advances
// all admitted runs together: every round issues a quantum of steps for
// each run, each on its own stream (one graph launch per step), so the
// runs' kernels overlap on the device; then it retires finished runs and
// admits queued ones while the HBM budget and the slot limit allow.  A run
// joins or leaves at a round boundary, so a late arrival does not wait for
// the runs already on the GPU to finish.  Admitted runs get their context
// (allocation, upload, init) on a helper thread, so a large run's creation
// does not hold up the quanta of the runs already on the GPU.
//
// A run's arithmetic never depends on what else is running: its result is
// the same as when it runs alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hd_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kResultFields[] = {HD_F_U, HD_F_V, HD_F_Z, HD_F_FS, HD_F_DH};
constexpr int kNumResults = 6;  // the five above + khouot
constexpr int kQuantumSpread = 4;  // most steps per round of a small run, in quanta

struct Run {
    int64_t id = 0;
    // owned copies of the inputs
    std::vector<double> grid[10];  // h, hsnham, VISCOIDX, Fw, Tsxw, Tsyw, u0, v0, z0, FS0
    std::vector<double> bc[4], CC[4];
    hd_mesh mesh{};
    hd_coeffs coeffs{};
    hd_options ops{};
    hd_config cfg{};
    size_t footprint = 0;
    int state = HD_RUN_QUEUED;
    double t = 0.0;
    int T = 0;  // Tmax truncated as engine.cpp:62
    long steps = 0;
    hd_ctx *ctx = nullptr;
    bool cancel = false;
    int rc = 0;
    std::string err;
    Clock::time_point submitted, started, finished;
    std::vector<double> results[kNumResults];
};

int result_slot(int field) {
    for (int k = 0; k < 5; k++)
        if (kResultFields[k] == field) return k;
    return field == HD_F_KHOUOT ? 5 : -1;
}

const double *own(std::vector<double> &dst, const double *src, size_t n) {
    if (!src) return nullptr;
    dst.assign(src, src + n);
    return dst.data();
}

}  // namespace

struct hd_server {
    hd_server_config cfg{};
    std::mutex mu;
    std::condition_variable wake, done;
    std::deque<Run *> queue;
    std::deque<Run *> to_create;  // admitted, waiting for the helper thread
    int creating = 0;             // admitted runs without a context yet
    std::vector<Run *> active;
    std::map<int64_t, std::unique_ptr<Run>> runs;
    size_t used = 0;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker, creator;
    std::condition_variable create_cv;
    // statistics
    std::vector<double> latency;  // seconds, completed runs
    double cell_steps = 0;
    long submitted = 0, admitted = 0, completed = 0, failed = 0, cancelled = 0;
    Clock::time_point first_submit;
    Clock::time_point last_finish;
    bool any = false;

    void loop();
    void create_loop();
    bool admissible() const;
    void finish(Run *r, int state);
};

// Whether the worker can act on the head of the queue now: a free slot and
// a head that fits the budget, or one that leaves the queue anyway.  Called
// with mu held.
bool hd_server::admissible() const {
    if (queue.empty() || (int)active.size() + creating >= cfg.max_active) return false;
    const Run *r = queue.front();
    return r->cancel || r->footprint > cfg.hbm_budget || used + r->footprint <= cfg.hbm_budget;
}

void hd_server::finish(Run *r, int state) {
    // called by the worker without the lock held
    if (r->ctx) {
        if (state == HD_RUN_DONE) {
            const size_t n = (size_t)(r->mesh.M + 3) * (r->mesh.N + 3);
            for (int k = 0; k < 5; k++) {
                r->results[k].resize(n);
                int rc = hd_copy_out(r->ctx, kResultFields[k], r->results[k].data(), n);
                if (rc && !r->rc) {
                    r->rc = rc;
                    r->err = hd_last_error(r->ctx);
                }
            }
            std::vector<int32_t> kh(n);
            int rc = hd_copy_out_i32(r->ctx, HD_F_KHOUOT, kh.data(), n);
            if (rc && !r->rc) {
                r->rc = rc;
                r->err = hd_last_error(r->ctx);
            }
            r->results[5].assign(kh.begin(), kh.end());
            if (r->rc) state = HD_RUN_FAILED;
        }
        hd_destroy(r->ctx);
        r->ctx = nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    r->state = state;
    r->finished = Clock::now();
    used -= r->footprint;
    if (state == HD_RUN_DONE) {
        completed++;
        latency.push_back(std::chrono::duration<double>(r->finished - r->submitted).count());
        last_finish = r->finished;
    } else if (state == HD_RUN_CANCELLED) {
        cancelled++;
    } else {
        failed++;
    }
    done.notify_all();
    wake.notify_all();  // budget freed: the queue's head may fit now
}

// The helper thread: creates and initialises the context of each admitted
// run in admission order, then hands the run to the worker.
void hd_server::create_loop() {
    (void)hipSetDevice(cfg.device);
    for (;;) {
        Run *r;
        {
            std::unique_lock<std::mutex> lk(mu);
            create_cv.wait(lk, [&] { return stop || !to_create.empty(); });
            if (stop) return;
            r = to_create.front();
            to_create.pop_front();
        }
        r->started = Clock::now();
        int rc = hd_create(&r->mesh, &r->coeffs, &r->ops, &r->cfg, &r->ctx);
        if (!rc) rc = hd_set_graph(r->ctx, cfg.use_graph);
        if (!rc) rc = hd_init(r->ctx);
        if (rc) {
            r->rc = rc;
            r->err = hd_last_error(r->ctx);
        }
        bool cancel;
        {
            std::lock_guard<std::mutex> lk(mu);
            creating--;
            cancel = r->cancel;
            if (!rc && !cancel) {
                r->state = HD_RUN_RUNNING;
                active.push_back(r);
            }
            wake.notify_all();
        }
        if (rc) finish(r, HD_RUN_FAILED);
        else if (cancel) finish(r, HD_RUN_CANCELLED);
    }
}

void hd_server::loop() {
    (void)hipSetDevice(cfg.device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            wake.wait(lk, [&] { return stop || !active.empty() || admissible(); });
            if (stop) return;
            // FIFO admission by HBM footprint and slot count; the contexts
            // are made on the helper thread
            while (!queue.empty() && (int)active.size() + creating < cfg.max_active) {
                Run *r = queue.front();
                if (r->cancel) {
                    queue.pop_front();
                    r->state = HD_RUN_CANCELLED;
                    cancelled++;
                    done.notify_all();
                    continue;
                }
                if (r->footprint > cfg.hbm_budget) {
                    queue.pop_front();
                    r->state = HD_RUN_FAILED;
                    r->rc = HD_E_NOMEM;
                    r->err = "run needs " + std::to_string(r->footprint) + " bytes, over the server's HBM budget";
                    failed++;
                    done.notify_all();
                    continue;
                }
                if (used + r->footprint > cfg.hbm_budget) break;
                queue.pop_front();
                used += r->footprint;
                admitted++;
                creating++;
                to_create.push_back(r);
                create_cv.notify_one();
            }
        }
        std::vector<Run *> cur;
        {
            std::lock_guard<std::mutex> lk(mu);
            cur = active;
        }
        // One quantum per run, each on its own stream.  The largest run
        // takes steps_per_quantum steps and sets the round's length; a
        // smaller run takes more steps in the same round, in inverse
        // proportion to its cell count (up to kQuantumSpread times as
        // many), so small runs finish and free their slots sooner instead
        // of advancing in lockstep with the largest one.
        double cells_max = 0;
        for (Run *r : cur) cells_max = std::max(cells_max, (double)r->mesh.M * r->mesh.N);
        for (Run *r : cur) {
            const double ratio = cells_max / ((double)r->mesh.M * r->mesh.N);
            const int quantum = cfg.steps_per_quantum * std::max(1, std::min(kQuantumSpread, (int)ratio));
            for (int k = 0; k < quantum && r->t < r->T && !r->rc; k++) {
                int rc = hd_step(r->ctx, &r->t);
                if (rc) {
                    r->rc = rc;
                    r->err = hd_last_error(r->ctx);
                } else {
                    r->steps++;
                }
            }
        }
        std::vector<std::pair<Run *, int>> retire;
        for (Run *r : cur) {
            if (!r->rc) {
                int rc = hd_sync(r->ctx);
                if (rc) {
                    r->rc = rc;
                    r->err = hd_last_error(r->ctx);
                }
            }
            bool cancel;
            {
                std::lock_guard<std::mutex> lk(mu);
                cancel = r->cancel;
            }
            if (r->rc)
                retire.push_back({r, HD_RUN_FAILED});
            else if (cancel)
                retire.push_back({r, HD_RUN_CANCELLED});
            else if (r->t >= r->T)
                retire.push_back({r, HD_RUN_DONE});
        }
        for (auto &rs : retire) {
            {
                std::lock_guard<std::mutex> lk(mu);
                active.erase(std::find(active.begin(), active.end(), rs.first));
                if (rs.second == HD_RUN_DONE)
                    cell_steps += (double)rs.first->mesh.M * rs.first->mesh.N * rs.first->steps;
            }
            finish(rs.first, rs.second);
        }
    }
}

extern "C" {

int hd_server_create(const hd_server_config *cfg, hd_server **out) {
    if (!out || !cfg) return HD_E_INVALID;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->steps_per_quantum < 1 || cfg->hbm_budget == 0) return HD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HD_E_NODEV;
    if (cfg->device < 0 || cfg->device >= ndev) return HD_E_NODEV;
    hd_server *s = new hd_server();
    s->cfg = *cfg;
    s->worker = std::thread([s] { s->loop(); });
    s->creator = std::thread([s] { s->create_loop(); });
    *out = s;
    return HD_OK;
}

int hd_server_submit(hd_server *s, const hd_mesh *mesh, const hd_coeffs *coeffs, const hd_options *ops,
                     const hd_config *cfg, int64_t *id) {
    if (!s || !mesh || !coeffs || !ops || !id) return HD_E_INVALID;
    if (mesh->M < 4 || mesh->N < 4 || !mesh->h || !mesh->hsnham || !mesh->VISCOIDX || mesh->total_time < 2)
        return HD_E_INVALID;
    for (int k = 0; k < 4; k++)
        if (!mesh->bc[k]) return HD_E_INVALID;
    auto r = std::make_unique<Run>();
    const size_t n = (size_t)(mesh->M + 3) * (mesh->N + 3), T = mesh->total_time;
    r->mesh = *mesh;
    r->mesh.h = own(r->grid[0], mesh->h, n);
    r->mesh.hsnham = own(r->grid[1], mesh->hsnham, n);
    r->mesh.VISCOIDX = own(r->grid[2], mesh->VISCOIDX, n);
    r->mesh.Fw = own(r->grid[3], mesh->Fw, n);
    r->mesh.Tsxw = own(r->grid[4], mesh->Tsxw, n);
    r->mesh.Tsyw = own(r->grid[5], mesh->Tsyw, n);
    r->mesh.u0 = own(r->grid[6], mesh->u0, n);
    r->mesh.v0 = own(r->grid[7], mesh->v0, n);
    r->mesh.z0 = own(r->grid[8], mesh->z0, n);
    r->mesh.FS0 = own(r->grid[9], mesh->FS0, n);
    for (int k = 0; k < 4; k++) {
        r->mesh.bc[k] = own(r->bc[k], mesh->bc[k], T);
        r->mesh.CC[k] = own(r->CC[k], mesh->CC[k], T);
    }
    r->coeffs = *coeffs;
    r->ops = *ops;
    if (cfg) r->cfg = *cfg;
    r->cfg.device = s->cfg.device;
    r->t = ops->t_start;
    r->T = (int)ops->Tmax;
    r->footprint = hd::footprint(&r->mesh, r->cfg.dense_layout, 1);
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->submitted = Clock::now();
    if (!s->any) {
        s->first_submit = r->submitted;
        s->any = true;
    }
    s->submitted++;
    *id = r->id;
    s->queue.push_back(r.get());
    s->runs.emplace(r->id, std::move(r));
    s->wake.notify_all();
    return HD_OK;
}

int hd_server_poll(hd_server *s, int64_t id, int *state, double *t) {
    if (!s) return HD_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->runs.find(id);
    if (it == s->runs.end()) return HD_E_INVALID;
    if (state) *state = it->second->state;
    if (t) *t = it->second->t;
    return HD_OK;
}

int hd_server_wait(hd_server *s, int64_t id, double timeout_s, int *state) {
    if (!s) return HD_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->runs.find(id);
    if (it == s->runs.end()) return HD_E_INVALID;
    Run *r = it->second.get();
    auto pred = [&] { return r->state >= HD_RUN_DONE; };
    if (timeout_s < 0)
        s->done.wait(lk, pred);
    else
        s->done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    if (state) *state = r->state;
    return r->state >= HD_RUN_DONE ? HD_OK : HD_E_STATE;
}

int hd_server_result(hd_server *s, int64_t id, int field, double *dst, size_t n) {
    if (!s || !dst) return HD_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->runs.find(id);
    if (it == s->runs.end()) return HD_E_INVALID;
    Run *r = it->second.get();
    if (r->state == HD_RUN_CANCELLED) return HD_E_CANCELLED;
    if (r->state != HD_RUN_DONE) return r->state == HD_RUN_FAILED ? (r->rc ? r->rc : HD_E_STATE) : HD_E_STATE;
    const int k = result_slot(field);
    if (k < 0 || n != r->results[k].size()) return HD_E_INVALID;
    std::memcpy(dst, r->results[k].data(), n * sizeof(double));
    return HD_OK;
}

int hd_server_error(hd_server *s, int64_t id, char *buf, size_t n) {
    if (!s || !buf || n == 0) return HD_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->runs.find(id);
    if (it == s->runs.end()) return HD_E_INVALID;
    std::snprintf(buf, n, "%s", it->second->err.c_str());
    return HD_OK;
}

int hd_server_cancel(hd_server *s, int64_t id) {
    if (!s) return HD_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->runs.find(id);
    if (it == s->runs.end()) return HD_E_INVALID;
    it->second->cancel = true;
    s->wake.notify_all();
    return HD_OK;
}

int hd_server_release(hd_server *s, int64_t id) {
    if (!s) return HD_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->runs.find(id);
    if (it == s->runs.end()) return HD_E_INVALID;
    if (it->second->state < HD_RUN_DONE) return HD_E_STATE;
    s->runs.erase(it);
    return HD_OK;
}

int hd_server_stats(hd_server *s, char *buf, size_t n) {
    if (!s || !buf || n == 0) return HD_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latency;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        size_t k = (size_t)std::ceil(q * l.size());
        return l[std::min(l.size(), std::max<size_t>(k, 1)) - 1];
    };
    double mean = 0;
    for (double x : l) mean += x;
    if (!l.empty()) mean /= l.size();
    const double wall = s->any ? std::chrono::duration<double>(Clock::now() - s->first_submit).count() : 0.0;
    const double busy = s->any && s->completed
                            ? std::chrono::duration<double>(s->last_finish - s->first_submit).count()
                            : 0.0;
    int written = std::snprintf(
        buf, n,
        "{\"submitted\": %ld, \"admitted\": %ld, \"completed\": %ld, \"failed\": %ld, \"cancelled\": %ld, "
        "\"queued\": %zu, \"active\": %zu, \"p50_s\": %.6f, \"p99_s\": %.6f, \"mean_s\": %.6f, "
        "\"cell_steps\": %.6e, \"wall_s\": %.6f, \"cell_steps_per_s\": %.6e, \"hbm_used\": %zu, "
        "\"hbm_budget\": %zu}",
        s->submitted, s->admitted, s->completed, s->failed, s->cancelled, s->queue.size(), s->active.size(),
        pct(0.5), pct(0.99), mean, s->cell_steps, wall, busy > 0 ? s->cell_steps / busy : 0.0, s->used,
        s->cfg.hbm_budget);
    return written > 0 && (size_t)written < n ? HD_OK : HD_E_INVALID;
}

int hd_server_destroy(hd_server *s) {
    if (!s) return HD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->wake.notify_all();
        s->create_cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    if (s->creator.joinable()) s->creator.join();
    // runs the helper thread had made a context for but not yet handed over
    // are in `active` too; runs still waiting for one have none
    for (Run *r : s->active)
        if (r->ctx) hd_destroy(r->ctx);
    delete s;
    return HD_OK;
}

}  // extern "C"
