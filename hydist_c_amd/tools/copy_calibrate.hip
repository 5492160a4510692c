// copy_calibrate - a plain device-to-device copy of a known size (256 MiB
// read + 256 MiB written per launch, three launches).  The FETCH_SIZE and
// WRITE_SIZE counters that rocprofv3 reports for it give the factor that
// turns those counters into HBM bytes for the library's kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

int main() {
    const size_t n = (size_t)32 << 20;  // 32 Mi doubles
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double)) != hipSuccess || hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 0, n * sizeof(double)) != hipSuccess) return 1;
    for (int r = 0; r < 3; r++) k_copy<<<8192, 256>>>(a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("k_copy: %zu KiB read and %zu KiB written per launch, 3 launches\n", n * sizeof(double) / 1024,
                n * sizeof(double) / 1024);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
