// hbm_copy_probe.hip -- a plain device-to-device copy of known size (512 MiB
// read + 512 MiB written per launch), run under rocprofv3 --pmc FETCH_SIZE /
// WRITE_SIZE to calibrate what those counters report per byte actually moved;
// the interpolation kernels' HBM traffic is scaled by the same factors.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e = (x);                                                        \
        if (e != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));            \
            return 1;                                                              \
        }                                                                          \
    } while (0)

int main() {
    const size_t bytes = size_t(512) << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int r = 0; r < 3; ++r) {
        CHECK(hipEventRecord(e0));
        k_copy<<<8192, 256>>>(a, b, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("copy of %zu MiB: %.3f ms, %.1f GB/s read + write\n", bytes >> 20, ms, 2.0 * bytes / ms / 1e6);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
