// Calibration program for HBM traffic counters: copies a buffer of known size with a plain vectorised kernel, so that
// rocprofv3's FETCH_SIZE / WRITE_SIZE of this dispatch (KiB) can be divided by the bytes it must move. Also prints the
// copy bandwidth it measures with HIP events.
//   hbm_copy [MiB per buffer, default 1024] [repeats, default 5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copyKernel(const float4 *__restrict__ in, float4 *__restrict__ out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e = (x);                                                         \
        if (e != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));            \
            return 1;                                                               \
        }                                                                           \
    } while (0)

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int grid = 256 * 32;
    hipLaunchKernelGGL(copyKernel, dim3(grid), dim3(256), 0, 0, a, b, n);   // warm-up
    CHECK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(copyKernel, dim3(grid), dim3(256), 0, 0, a, b, n);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    std::printf("hbm_copy: %zu MiB read + %zu MiB written per dispatch, best %.3f ms, %.1f GB/s\n", mib, mib, best,
                2.0 * bytes / (best * 1e-3) / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
