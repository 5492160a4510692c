// CPU emulation of the subset of the HIP runtime and device language that
// libhydist_amd uses, for bounds-checking its kernels on the host.
//
// Test infrastructure only: the library built against this header runs
// every kernel on the CPU (each workgroup's lanes as fibers, __syncthreads
// a yield point) with device allocations placed against guard pages, so an
// out-of-bounds access stops at the faulting kernel and lane.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__ __restrict
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) alignas(n)

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// 16-byte vector of the device code's paired stores; a store through a
// double2 pointer that is not 16-byte aligned aborts with a message
struct alignas(16) double2 {
    double x, y;
    double2() = default;
    double2(double a, double b) : x(a), y(b) {}
    double2(const double2 &) = default;
    double2 &operator=(const double2 &o) {
        if (reinterpret_cast<uintptr_t>(this) & 15) {
            std::fprintf(stderr, "emu: misaligned double2 store at %p\n", (void *)this);
            std::abort();
        }
        x = o.x;
        y = o.y;
        return *this;
    }
};
inline double2 make_double2(double x, double y) { return double2(x, y); }
struct alignas(8) int2 {
    int x, y;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNoDevice = 100 };
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph *hipGraphExec_t;
typedef struct emu_graph_node *hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
enum { hipStreamDefault = 0, hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
enum hipMemoryType { hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2 };
struct hipPointerAttribute_t {
    hipMemoryType type;
    int device;
    void *devicePointer, *hostPointer;
};
struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    int l2CacheSize;
    size_t totalGlobalMem;
};

// runtime
hipError_t hipGetDeviceCount(int *n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipMalloc(void **p, size_t n);
template <class T>
inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
hipError_t hipFree(void *p);
hipError_t hipMallocAsync(void **p, size_t n, hipStream_t s);
hipError_t hipFreeAsync(void *p, hipStream_t s);
hipError_t hipGetDevice(int *d);
// A device symbol's address: a guarded buffer of the symbol's size, one per
// symbol (kernels reach the symbol only through the pointer returned here,
// so accesses past it fault like those past any device buffer).
void *emu_symbol_buffer(const void *sym, size_t n);
template <class T>
inline hipError_t hipGetSymbolAddress(void **p, T *sym) {
    *p = emu_symbol_buffer((const void *)sym, sizeof(T));
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipMemset(void *p, int v, size_t n);
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t s = nullptr);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemcpy2D(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind k);
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind k,
                            hipStream_t st = nullptr);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
enum { hipEventDefault = 0, hipEventDisableTiming = 2 };
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute a, int v);
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *p);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *n, char *log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);

// device side
namespace hd_emu {
struct Lane {
    dim3 tid;
    int linear;
};
Lane *cur();
extern dim3 g_block_idx, g_block_dim, g_grid_dim;
void sync();
unsigned long long ballot(bool p);
void *dyn_lds();
void launch(const char *name, dim3 g, dim3 b, size_t shmem, hipStream_t st, std::function<void()> body);

template <class K>
struct Launcher {
    const char *name;
    K k;
    dim3 g, b;
    size_t sm;
    hipStream_t st;
    template <class... A>
    void operator()(A... a) const {
        K kk = k;
        launch(name, g, b, sm, st, [=]() { kk(a...); });
    }
};
template <class K>
Launcher<K> launcher(const char *name, K k, dim3 g, dim3 b, size_t sm = 0, hipStream_t st = nullptr) {
    return Launcher<K>{name, k, g, b, sm, st};
}
}  // namespace hd_emu

#define threadIdx (hd_emu::cur()->tid)
#define blockIdx (hd_emu::g_block_idx)
#define blockDim (hd_emu::g_block_dim)
#define gridDim (hd_emu::g_grid_dim)

inline void __syncthreads() { hd_emu::sync(); }
inline unsigned long long __ballot(int p) { return hd_emu::ballot(p != 0); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __double2int_rn(double x) { return (int)std::nearbyint(x); }
template <class T>
inline T atomicOr(T *p, T v) {
    T o = *p;
    *p = o | v;
    return o;
}
inline int atomicAdd(int *p, int v) {
    int o = *p;
    *p = o + v;
    return o;
}
inline double atomicAdd(double *p, double v) {
    double o = *p;
    *p = o + v;
    return o;
}
using std::isfinite;
using std::isnan;
using std::max;
using std::min;
