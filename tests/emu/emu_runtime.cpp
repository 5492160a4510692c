// CPU runtime behind tests/emu/hip/hip_runtime.h (test infrastructure).
//
// Device allocations end against a PROT_NONE guard page (another guards the
// start), so a read or write past a buffer's end faults at once; the
// SIGSEGV handler names the kernel, workgroup and lane that faulted and the
// allocation the address belongs to.  A launch runs its workgroups in
// order; a workgroup's lanes are ucontext fibers and __syncthreads is a
// yield: a pass resumes every live lane once, so no lane leaves a barrier
// before every lane has reached it.  Launches and allocations from several
// host threads (the serving loop's worker and its context-making helper)
// take one lock, so they run one after another, as on one queue.  Graph
// capture is per stream, as in HIP's relaxed capture mode: while one
// thread captures its stream, launches on other streams run at once.
// Built with -fsanitize=thread, the fibers are announced to
// ThreadSanitizer, which then orders a workgroup's lanes by their switches.
// HD_EMU_ORDER=reverse runs the workgroups of a launch, and the lanes of
// each pass between barriers, in reverse order; HD_EMU_ORDER=shuffle:SEED
// in a pseudo-random order drawn anew for every launch and pass.  A result
// that changes with either depends on an order the hardware does not
// promise (a missing barrier, or one workgroup reading what another writes
// in the same launch).
#include <signal.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hip/hip_runtime.h"

#if defined(__SANITIZE_THREAD__)
#include <sanitizer/tsan_interface.h>
#define EMU_TSAN 1
#else
#define EMU_TSAN 0
#endif

namespace {

struct Alloc {
    char *base;      // mapping start (guard page)
    size_t total;    // mapping length
    size_t n;        // requested bytes
};
std::map<uintptr_t, Alloc> g_allocs;  // by user pointer
size_t g_page = 4096;
std::recursive_mutex g_mu;  // launches and the allocation map

void *guarded(size_t n, Alloc *out) {
    const size_t body = (n + 15) / 16 * 16;
    const size_t pages = (body + g_page - 1) / g_page;
    const size_t total = (pages + 2) * g_page;
    char *base = (char *)mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (base == MAP_FAILED) return nullptr;
    mprotect(base, g_page, PROT_NONE);
    mprotect(base + total - g_page, g_page, PROT_NONE);
    char *user = base + total - g_page - body;
    *out = Alloc{base, total, n};
    return user;
}

// --- launch state ---
const char *g_kernel = "(host)";
struct Fiber {
    ucontext_t ctx;
    hd_emu::Lane lane;
    bool done;
    char *stack;
    void *tsan;  // the fiber's ThreadSanitizer identity (sanitized builds)
};
constexpr size_t kStack = 256 * 1024;
std::vector<Fiber> g_fibers;
ucontext_t g_sched;
void *g_sched_tsan = nullptr;  // the launching thread's ThreadSanitizer identity
int g_cur = -1;
const std::function<void()> *g_body = nullptr;
void *g_dyn = nullptr;
unsigned long long g_ballot[2][16];
int g_ballot_parity[1024];

// Announces the switch to a fiber (or back to the launching thread) to
// ThreadSanitizer; flag 0 orders what ran before the switch before what
// runs after it.  Always inlined, so it adds no frame of its own to pop
// after the switch.
inline __attribute__((always_inline)) void tsan_switch(void *to) {
#if EMU_TSAN
    __tsan_switch_to_fiber(to, 0);
#else
    (void)to;
#endif
}

// Not instrumented: its return, after the switch back to the launching
// thread, must not pop that thread's ThreadSanitizer frames.
__attribute__((no_sanitize_thread)) void fiber_entry() {
    (*g_body)();
    g_fibers[g_cur].done = true;
    tsan_switch(g_sched_tsan);  // returns to g_sched through uc_link
}

bool g_capturing = false;
hipStream_t g_capture_stream = nullptr;
bool g_reverse = false;  // HD_EMU_ORDER=reverse
bool g_shuffle = false;  // HD_EMU_ORDER=shuffle:SEED
std::mt19937_64 g_rng;
struct Node {
    const char *name;
    dim3 g, b;
    size_t sm;
    std::function<void()> body;
};
std::vector<Node> *g_capture = nullptr;

void run_launch(const char *name, dim3 g, dim3 b, size_t sm, const std::function<void()> &body) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    const unsigned nt = b.x * b.y * b.z;
    if (nt == 0 || nt > 1024) {
        std::fprintf(stderr, "[emu] %s: bad block size %u\n", name, nt);
        std::abort();
    }
    if (g.x == 0 || g.y == 0 || g.z == 0 || g.x > 0x7fffffffu || g.y > 65535 || g.z > 65535) {
        std::fprintf(stderr, "[emu] %s: grid (%u, %u, %u) outside the device's limits\n", name, g.x, g.y, g.z);
        std::abort();
    }
    if (sm > 160 * 1024) {
        std::fprintf(stderr, "[emu] %s: dynamic LDS %zu > 160 KiB\n", name, sm);
        std::abort();
    }
    Alloc lds{};
    g_dyn = sm ? guarded(sm, &lds) : nullptr;
    g_kernel = name;
    hd_emu::g_grid_dim = g;
    hd_emu::g_block_dim = b;
    if (g_fibers.size() < nt) {
        size_t old = g_fibers.size();
        g_fibers.resize(nt);
        for (size_t i = old; i < nt; i++) {
            g_fibers[i].stack = (char *)mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS,
                                             -1, 0);
#if EMU_TSAN
            g_fibers[i].tsan = __tsan_create_fiber(0);
#endif
        }
    }
#if EMU_TSAN
    g_sched_tsan = __tsan_get_current_fiber();
#endif
    g_body = &body;
    const size_t nblocks = (size_t)g.x * g.y * g.z;
    std::vector<size_t> border;
    std::vector<unsigned> lorder(nt);
    std::iota(lorder.begin(), lorder.end(), 0u);
    if (g_reverse) std::reverse(lorder.begin(), lorder.end());
    if (g_shuffle) {
        border.resize(nblocks);
        std::iota(border.begin(), border.end(), (size_t)0);
        std::shuffle(border.begin(), border.end(), g_rng);
    }
    for (size_t q = 0; q < nblocks; q++) {
        const size_t blk = g_shuffle ? border[q] : (g_reverse ? nblocks - 1 - q : q);
        hd_emu::g_block_idx = dim3((unsigned)(blk % g.x), (unsigned)(blk / g.x % g.y),
                                   (unsigned)(blk / ((size_t)g.x * g.y)));
        if (sm) std::memset(g_dyn, 0xcd, sm);
        for (unsigned t = 0; t < nt; t++) {
            Fiber &f = g_fibers[t];
            f.lane.tid = dim3(t % b.x, (t / b.x) % b.y, t / (b.x * b.y));
            f.lane.linear = (int)t;
            f.done = false;
            g_ballot_parity[t] = 0;
            getcontext(&f.ctx);
            f.ctx.uc_stack.ss_sp = f.stack;
            f.ctx.uc_stack.ss_size = kStack;
            f.ctx.uc_link = &g_sched;
            makecontext(&f.ctx, fiber_entry, 0);
        }
        std::memset(g_ballot, 0, sizeof(g_ballot));
        bool alive = true;
        while (alive) {
            alive = false;
            if (g_shuffle) std::shuffle(lorder.begin(), lorder.end(), g_rng);
            for (unsigned k = 0; k < nt; k++) {
                const unsigned t = lorder[k];
                if (g_fibers[t].done) continue;
                g_cur = (int)t;
                tsan_switch(g_fibers[t].tsan);
                swapcontext(&g_sched, &g_fibers[t].ctx);
                if (!g_fibers[t].done) alive = true;
            }
        }
        g_cur = -1;
    }
    if (sm) munmap(lds.base, lds.total);
    g_dyn = nullptr;
    g_kernel = "(host)";
}

void segv(int, siginfo_t *si, void *) {
    const uintptr_t a = (uintptr_t)si->si_addr;
    char buf[512];
    int n = std::snprintf(buf, sizeof(buf), "[emu] FAULT at %p in %s", si->si_addr, g_kernel);
    if (g_cur >= 0) {
        const hd_emu::Lane &l = g_fibers[g_cur].lane;
        n += std::snprintf(buf + n, sizeof(buf) - n, " block (%u,%u,%u) thread (%u,%u,%u)", hd_emu::g_block_idx.x,
                           hd_emu::g_block_idx.y, hd_emu::g_block_idx.z, l.tid.x, l.tid.y, l.tid.z);
    }
    for (auto &kv : g_allocs) {
        const Alloc &al = kv.second;
        if (a >= (uintptr_t)al.base && a < (uintptr_t)al.base + al.total) {
            n += std::snprintf(buf + n, sizeof(buf) - n, " -- allocation %p of %zu bytes, offset %lld",
                               (void *)kv.first, al.n, (long long)(a - kv.first));
            break;
        }
    }
    n += std::snprintf(buf + n, sizeof(buf) - n, "\n");
    (void)!write(2, buf, n);
    _exit(86);
}

struct Init {
    Init() {
        g_page = (size_t)sysconf(_SC_PAGESIZE);
        static char altstack[1 << 16];
        stack_t ss{};
        ss.ss_sp = altstack;
        ss.ss_size = sizeof(altstack);
        sigaltstack(&ss, nullptr);
        struct sigaction sa{};
        sa.sa_sigaction = segv;
        sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
        sigaction(SIGSEGV, &sa, nullptr);
        sigaction(SIGBUS, &sa, nullptr);
        const char *order = std::getenv("HD_EMU_ORDER");
        g_reverse = order && std::string(order) == "reverse";
        if (order && std::string(order).rfind("shuffle:", 0) == 0) {
            g_shuffle = true;
            g_rng.seed(std::strtoull(order + 8, nullptr, 10));
        }
    }
} g_init;

}  // namespace

namespace hd_emu {
dim3 g_block_idx, g_block_dim, g_grid_dim;
Lane *cur() { return &g_fibers[g_cur].lane; }
void sync() {
    tsan_switch(g_sched_tsan);
    swapcontext(&g_fibers[g_cur].ctx, &g_sched);
}
unsigned long long ballot(bool p) {
    const int t = g_cur, w = t / 64, lane = t % 64;
    int &par = g_ballot_parity[t];
    unsigned long long *acc = &g_ballot[par][w];
    if (p) *acc |= 1ull << lane;
    sync();
    const unsigned long long r = *acc;
    sync();
    if (lane == 0) *acc = 0;
    par ^= 1;
    return r;
}
void *dyn_lds() { return g_dyn; }
void launch(const char *name, dim3 g, dim3 b, size_t sm, hipStream_t st, std::function<void()> body) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (g_capturing && st == g_capture_stream) {
        g_capture->push_back(Node{name, g, b, sm, std::move(body)});
        return;
    }
    run_launch(name, g, b, sm, body);
}
}  // namespace hd_emu

struct emu_graph {
    std::vector<Node> nodes;
};

hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "CPU emulation");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "emu");
    p->multiProcessorCount = 1;
    p->totalGlobalMem = (size_t)8 << 30;
    return hipSuccess;
}
hipError_t hipMalloc(void **p, size_t n) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Alloc a{};
    void *u = guarded(n ? n : 1, &a);
    if (!u) return hipErrorOutOfMemory;
    g_allocs[(uintptr_t)u] = a;
    *p = u;
    return hipSuccess;
}
hipError_t hipFree(void *p) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    auto it = g_allocs.find((uintptr_t)p);
    if (it == g_allocs.end()) return p ? hipErrorInvalidValue : hipSuccess;
    munmap(it->second.base, it->second.total);
    g_allocs.erase(it);
    return hipSuccess;
}
hipError_t hipMallocAsync(void **p, size_t n, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipFreeAsync(void *p, hipStream_t) { return hipFree(p); }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
void *emu_symbol_buffer(const void *sym, size_t n) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    static std::map<const void *, void *> bufs;
    auto it = bufs.find(sym);
    if (it != bufs.end()) return it->second;
    void *p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) return nullptr;
    std::memset(p, 0, n);
    bufs[sym] = p;
    return p;
}
// "Device" buffers for tests that call entry points taking device pointers.
extern "C" void *hd_emu_alloc(size_t n) {
    void *p = nullptr;
    return hipMalloc(&p, n) == hipSuccess ? p : nullptr;
}
extern "C" void hd_emu_free(void *p) { (void)hipFree(p); }

// A call HIP would refuse (or record wrongly) on a capturing stream.
static void no_capture(const char *what, hipStream_t st) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (g_capturing && st == g_capture_stream) {
        std::fprintf(stderr, "[emu] %s during graph capture\n", what);
        std::abort();
    }
}
hipError_t hipMemset(void *p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t st) {
    no_capture("hipMemsetAsync", st);
    std::memset(p, v, n);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st) {
    no_capture("hipMemcpyAsync", st);
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpy2D(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind) {
    for (size_t r = 0; r < h; r++) std::memmove((char *)d + r * dp, (const char *)s + r * sp, w);
    return hipSuccess;
}
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind k,
                            hipStream_t st) {
    no_capture("hipMemcpy2DAsync", st);
    return hipMemcpy2D(d, dp, s, sp, w, h, k);
}
struct emu_stream {
    int dummy;
};
struct emu_event {
    int dummy;
};
hipError_t hipStreamCreate(hipStream_t *s) { *s = new emu_stream{0}; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t st) {
    no_capture("hipStreamSynchronize", st);
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event{0}; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
// Streams run in issue order here, so ordering events are no-ops (also
// under capture, where launches from every stream are recorded in issue
// order).
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int v) {
    return v <= 160 * 1024 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *p) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    std::memset(a, 0, sizeof(*a));
    for (auto &kv : g_allocs)
        if ((uintptr_t)p >= kv.first && (uintptr_t)p < kv.first + kv.second.n) {
            a->type = hipMemoryTypeDevice;
            a->devicePointer = (void *)p;
            return hipSuccess;
        }
    a->type = hipMemoryTypeHost;
    a->hostPointer = (void *)p;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (g_capturing) return hipErrorInvalidValue;  // one capture at a time here
    g_capturing = true;
    g_capture_stream = st;
    g_capture = new std::vector<Node>();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t *g) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g_capturing || st != g_capture_stream) return hipErrorInvalidValue;
    g_capturing = false;
    g_capture_stream = nullptr;
    *g = new emu_graph{std::move(*g_capture)};
    delete g_capture;
    g_capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    *x = new emu_graph{g->nodes};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    for (auto &nd : x->nodes) run_launch(nd.name, nd.g, nd.b, nd.sm, nd.body);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t x) { delete x; return hipSuccess; }
