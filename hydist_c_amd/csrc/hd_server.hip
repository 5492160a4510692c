// hd_server.hip - serving loop for what-if runs on one GPU.
//
// The reference runs one scenario per process, start to finish
// (main.cpp:138-309).  The server instead takes runs as they arrive
// (different meshes, hydrographs, roughness, durations, forms), keeps each
// admitted run's state resident in HBM, and advances all admitted runs
// together: every round issues a quantum of steps for each group of runs,
// each group on its own stream (one graph launch per step), so the groups'
// kernels overlap on the device; then it retires finished groups and
// admits queued runs while the HBM budget and the slot limit allow.  A run
// joins or leaves at a round boundary, so a late arrival does not wait for
// the runs already on the GPU to finish.
//
// A group is one context.  Queued runs that can share launches (the same
// mesh size, duration, form, options and coefficients; their inputs, such
// as depths, roughness and boundary series, may differ) are admitted
// together into one batched context (hd_batch_create), so every kernel
// launch of a step serves all of them (the scenario is the grid's outer
// dimension); other runs get a context of their own.  Contexts are made on
// a helper thread, so a large run's allocation and upload do not hold up
// the quanta of the groups already on the GPU.
//
// A run's arithmetic never depends on what else is running: its result is
// the same as when it runs alone (a batched context computes each scenario
// as a context of its own would).
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

constexpr int kResultFields[] = {HD_F_U, HD_F_V, HD_F_Z, HD_F_FS, HD_F_DH, HD_F_KHOUOT};
constexpr int kNumResults = 6;
constexpr int kQuantumSpread = 4;  // most steps per round of a small group, in quanta
constexpr int kMaxBatch = 16;      // runs per batched context at most

struct Group;

struct Run {
    int64_t id = 0;
    // owned copies of the inputs
    std::vector<double> grid[10];  // h, hsnham, VISCOIDX, Fw, Tsxw, Tsyw, u0, v0, z0, FS0
    std::vector<double> bc[4], CC[4];
    hd_mesh mesh{};
    hd_coeffs coeffs{};
    hd_options ops{};
    hd_config cfg{};
    size_t footprint = 0;  // of the run in a context of its own
    int state = HD_RUN_QUEUED;
    double t = 0.0;
    int T = 0;  // Tmax truncated as engine.cpp:62
    long steps = 0;
    Group *group = nullptr;
    int scen = 0;  // the run's scenario in its group's context
    bool cancel = false;
    int rc = 0;
    std::string err;
    Clock::time_point submitted;
    std::vector<double> results[kNumResults];
};

// Runs advanced by one context: scenario k is members[k].
struct Group {
    std::vector<Run *> members;
    hd_ctx *ctx = nullptr;
    double t = 0.0;
    int T = 0;
    long steps = 0;
    size_t footprint = 0;
    int rc = 0;
    std::string err;
    double cells() const { return (double)members[0]->mesh.M * members[0]->mesh.N * members.size(); }
};

int result_slot(int field) {
    for (int k = 0; k < kNumResults; k++)
        if (kResultFields[k] == field) return k;
    return -1;
}

const double *own(std::vector<double> &dst, const double *src, size_t n) {
    if (!src) return nullptr;
    dst.assign(src, src + n);
    return dst.data();
}

// Whether two runs can share one batched context: one mesh size and series
// length, one duration and start, one form and configuration, the same
// coefficients and the same options (hd_batch_create takes one of each).
bool same_launches(const Run &a, const Run &b) {
    const hd_options &x = a.ops, &y = b.ops;
    return a.mesh.M == b.mesh.M && a.mesh.N == b.mesh.N && a.mesh.total_time == b.mesh.total_time && a.T == b.T &&
           a.t == b.t && std::memcmp(&a.cfg, &b.cfg, sizeof(hd_config)) == 0 &&
           std::memcmp(&a.coeffs, &b.coeffs, sizeof(hd_coeffs)) == 0 && x.Tmax == y.Tmax &&
           x.t_start == y.t_start && x.sediment_start == y.sediment_start &&
           x.bed_change_start == y.bed_change_start && x.interval == y.interval && x.kenhhepd == y.kenhhepd &&
           x.kenhhepng == y.kenhhepng && x.ketdinh == y.ketdinh && x.channel == y.channel && x.debug == y.debug &&
           x.plot == y.plot && x.total_time == y.total_time;
}

}  // namespace

struct hd_server {
    hd_server_config cfg{};
    std::mutex mu;
    std::condition_variable wake, done, create_cv;
    std::deque<Run *> queue;
    std::deque<Group *> to_create;  // admitted, waiting for the helper thread
    std::vector<Group *> active;
    int slots = 0;  // admitted runs whose group has not retired
    std::map<int64_t, std::unique_ptr<Run>> runs;
    size_t used = 0;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker, creator;
    // statistics
    std::vector<double> latency;  // seconds, completed runs
    double cell_steps = 0;
    long submitted = 0, admitted = 0, completed = 0, failed = 0, cancelled = 0;
    long batched = 0;  // runs admitted into a group of two or more
    Clock::time_point first_submit;
    Clock::time_point last_finish;
    bool any = false;

    void loop();
    void create_loop();
    bool admissible() const;
    void admit();
    void retire(Group *g, int state);
};

// Whether the worker can act on the head of the queue now: a free slot and
// a head that fits the budget, or one that leaves the queue anyway.  Called
// with mu held.
bool hd_server::admissible() const {
    if (queue.empty() || slots >= cfg.max_active) return false;
    const Run *r = queue.front();
    return r->cancel || r->footprint > cfg.hbm_budget || used + r->footprint <= cfg.hbm_budget;
}

// FIFO admission by HBM footprint and slot count: the queue's head, with
// the queued runs that can share its launches, becomes a group for the
// helper thread.  Called with mu held.
void hd_server::admit() {
    while (!queue.empty() && slots < cfg.max_active) {
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
        std::vector<Run *> m{r};
        for (size_t q = 1; q < queue.size() && (int)m.size() < kMaxBatch && slots + (int)m.size() < cfg.max_active;
             q++)
            if (!queue[q]->cancel && same_launches(*r, *queue[q])) m.push_back(queue[q]);
        size_t fp = m.size() == 1 ? r->footprint : hd::footprint(&r->mesh, r->cfg.dense_layout, (int)m.size());
        while (m.size() > 1 && used + fp > cfg.hbm_budget) {
            m.pop_back();
            fp = m.size() == 1 ? r->footprint : hd::footprint(&r->mesh, r->cfg.dense_layout, (int)m.size());
        }
        for (Run *o : m) queue.erase(std::find(queue.begin(), queue.end(), o));
        Group *g = new Group();
        g->members = m;
        g->footprint = fp;
        g->t = r->t;
        g->T = r->T;
        for (size_t k = 0; k < m.size(); k++) {
            m[k]->group = g;
            m[k]->scen = (int)k;
        }
        used += fp;
        slots += (int)m.size();
        admitted += (long)m.size();
        if (m.size() > 1) batched += (long)m.size();
        to_create.push_back(g);
        create_cv.notify_one();
    }
}

// Ends a group: copies its runs' results out when it completed, frees its
// context and its budget, and records each run's end.  Called without mu
// held, by the worker or the helper thread; the group is deleted.  The
// copies wait for the device, so they go into the group's own buffers
// without mu and reach the runs (which API calls read under mu) with it.
void hd_server::retire(Group *g, int state) {
    const size_t nm = g->members.size();
    std::vector<char> skip(nm);
    {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t m = 0; m < nm; m++) skip[m] = g->members[m]->cancel;
    }
    std::vector<std::vector<double>> out(nm * kNumResults);
    std::vector<int> rcs(nm, 0);
    std::vector<std::string> errs(nm);
    if (g->ctx && state == HD_RUN_DONE) {
        const size_t n = (size_t)(g->members[0]->mesh.M + 3) * (g->members[0]->mesh.N + 3);
        for (size_t m = 0; m < nm; m++) {
            if (skip[m]) continue;
            for (int k = 0; k < kNumResults && !rcs[m]; k++) {
                std::vector<double> &dst = out[m * kNumResults + k];
                dst.resize(n);
                int rc = hd_batch_copy_out(g->ctx, g->members[m]->scen, kResultFields[k], dst.data(), n);
                if (rc) {
                    rcs[m] = rc;
                    errs[m] = hd_last_error(g->ctx);
                }
            }
        }
    }
    if (g->ctx) hd_destroy(g->ctx);
    g->ctx = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    const Clock::time_point now = Clock::now();
    for (size_t m = 0; m < nm; m++) {
        Run *r = g->members[m];
        r->t = g->t;
        r->steps = g->steps;
        for (int k = 0; k < kNumResults; k++) r->results[k] = std::move(out[m * kNumResults + k]);
        if (rcs[m]) {
            r->rc = rcs[m];
            r->err = errs[m];
        } else if (g->rc) {
            r->rc = g->rc;
            r->err = g->err;
        }
        int st = state;
        if (r->cancel) st = HD_RUN_CANCELLED;
        else if (st == HD_RUN_DONE && r->rc) st = HD_RUN_FAILED;
        r->state = st;
        r->group = nullptr;
        if (st == HD_RUN_DONE) {
            completed++;
            latency.push_back(std::chrono::duration<double>(now - r->submitted).count());
            last_finish = now;
            cell_steps += (double)r->mesh.M * r->mesh.N * r->steps;
        } else if (st == HD_RUN_CANCELLED) {
            cancelled++;
        } else {
            failed++;
        }
    }
    used -= g->footprint;
    slots -= (int)g->members.size();
    auto it = std::find(active.begin(), active.end(), g);
    if (it != active.end()) active.erase(it);
    delete g;
    done.notify_all();
    wake.notify_all();  // budget and slots freed: the queue's head may fit now
}

// The helper thread: creates and initialises the context of each admitted
// group in admission order, then hands the group to the worker.
void hd_server::create_loop() {
    (void)hipSetDevice(cfg.device);
    for (;;) {
        Group *g;
        {
            std::unique_lock<std::mutex> lk(mu);
            create_cv.wait(lk, [&] { return stop || !to_create.empty(); });
            if (stop) return;
            g = to_create.front();
            to_create.pop_front();
        }
        std::vector<hd_mesh> meshes;
        for (Run *r : g->members) meshes.push_back(r->mesh);
        const Run *r0 = g->members[0];
        int rc = hd_batch_create(meshes.data(), (int)meshes.size(), &r0->coeffs, &r0->ops, &r0->cfg, &g->ctx);
        if (!rc) rc = hd_set_graph(g->ctx, cfg.use_graph);
        if (!rc) rc = hd_init(g->ctx);
        if (rc) {
            g->rc = rc;
            g->err = hd_last_error(g->ctx);
        }
        bool all_cancelled = true;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Run *r : g->members) all_cancelled = all_cancelled && r->cancel;
            if (!rc && !all_cancelled) {
                for (Run *r : g->members) r->state = HD_RUN_RUNNING;
                active.push_back(g);
            }
            wake.notify_all();
        }
        if (rc) retire(g, HD_RUN_FAILED);
        else if (all_cancelled) retire(g, HD_RUN_CANCELLED);
    }
}

void hd_server::loop() {
    (void)hipSetDevice(cfg.device);
    for (;;) {
        std::vector<Group *> cur;
        {
            std::unique_lock<std::mutex> lk(mu);
            wake.wait(lk, [&] { return stop || !active.empty() || admissible(); });
            if (stop) return;
            admit();
            cur = active;
        }
        // One quantum per group, each on its own stream.  The largest group
        // (by cells over its runs) takes steps_per_quantum steps and sets
        // the round's length; a smaller group takes more steps in the same
        // round, in inverse proportion to its cell count (up to
        // kQuantumSpread times as many), so small runs finish and free their
        // slots sooner instead of advancing in lockstep with the largest.
        double cells_max = 0;
        for (Group *g : cur) cells_max = std::max(cells_max, g->cells());
        for (Group *g : cur) {
            const int quantum =
                cfg.steps_per_quantum * std::max(1, std::min(kQuantumSpread, (int)(cells_max / g->cells())));
            for (int k = 0; k < quantum && g->t < g->T && !g->rc; k++) {
                int rc = hd_step(g->ctx, &g->t);
                if (rc) {
                    g->rc = rc;
                    g->err = hd_last_error(g->ctx);
                } else {
                    g->steps++;
                }
            }
        }
        std::vector<std::pair<Group *, int>> ending;
        for (Group *g : cur) {
            if (!g->rc) {
                int rc = hd_sync(g->ctx);
                if (rc) {
                    g->rc = rc;
                    g->err = hd_last_error(g->ctx);
                }
            }
            bool all_cancelled = true;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Run *r : g->members) {
                    r->t = g->t;
                    all_cancelled = all_cancelled && r->cancel;
                }
            }
            if (g->rc)
                ending.push_back({g, HD_RUN_FAILED});
            else if (all_cancelled)
                ending.push_back({g, HD_RUN_CANCELLED});
            else if (g->t >= g->T)
                ending.push_back({g, HD_RUN_DONE});
        }
        for (auto &e : ending) retire(e.first, e.second);
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
    if (hd::check_setup(coeffs, cfg)) return HD_E_INVALID;  // refused now, not when its context is made
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
    Run *r = it->second.get();
    r->cancel = true;
    // a run still in the queue (not yet in a group) leaves it at once, even
    // while every slot is taken; one in a group ends at its group's next
    // round boundary
    if (r->state == HD_RUN_QUEUED && !r->group) {
        auto q = std::find(s->queue.begin(), s->queue.end(), r);
        if (q != s->queue.end()) {
            s->queue.erase(q);
            r->state = HD_RUN_CANCELLED;
            s->cancelled++;
            s->done.notify_all();
        }
    }
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
        "\"queued\": %zu, \"active\": %d, \"batched\": %ld, \"p50_s\": %.6f, \"p99_s\": %.6f, \"mean_s\": %.6f, "
        "\"cell_steps\": %.6e, \"wall_s\": %.6f, \"cell_steps_per_s\": %.6e, \"hbm_used\": %zu, "
        "\"hbm_budget\": %zu}",
        s->submitted, s->admitted, s->completed, s->failed, s->cancelled, s->queue.size(), s->slots, s->batched,
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
    // groups the helper thread made a context for are in `active`; groups
    // still waiting for one have none
    for (Group *g : s->active) {
        if (g->ctx) hd_destroy(g->ctx);
        delete g;
    }
    for (Group *g : s->to_create) delete g;
    delete s;
    return HD_OK;
}

}  // extern "C"
