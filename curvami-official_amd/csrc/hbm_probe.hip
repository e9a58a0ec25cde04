// HBM calibration probe for the profiler's FETCH_SIZE / WRITE_SIZE counters:
// copies `mib` MiB (default 1024) from one device buffer to another with a plain
// vector-load/vector-store kernel, `reps` times. Each dispatch moves exactly
// mib MiB in and mib MiB out, so the counters' per-dispatch readings divided by
// that size give the scale factor used to turn counter values into bytes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t bytes = mib << 20, n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    auto ok = [](hipError_t e, const char* what) {
        if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        return e == hipSuccess;
    };
    if (!ok(hipMemset(a, 0, bytes), "hipMemset")) return 1;
    hipEvent_t e0, e1;
    if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        if (!ok(hipEventRecord(e0), "hipEventRecord")) return 1;
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
        if (!ok(hipGetLastError(), "k_copy launch") || !ok(hipEventRecord(e1), "hipEventRecord") ||
            !ok(hipEventSynchronize(e1), "hipEventSynchronize"))
            return 1;
        float ms = 0;
        if (!ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) return 1;
        if (ms < best) best = ms;
    }
    std::printf("hbm_probe: %zu MiB copy, best %.3f ms, %.1f GB/s (read + write)\n", mib, best,
                2.0 * bytes / (best * 1e-3) / 1e9);
    ok(hipEventDestroy(e0), "hipEventDestroy");
    ok(hipEventDestroy(e1), "hipEventDestroy");
    ok(hipFree(a), "hipFree");
    ok(hipFree(b), "hipFree");
    return 0;
}
