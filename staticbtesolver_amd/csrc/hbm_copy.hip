// hbm_copy.hip -- calibration probe for HBM byte counters: copies a buffer of known size with a
// plain vectorised kernel so that rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch can be compared
// with the exact number of bytes moved.  Usage: hbm_copy [MiB] [launches]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4* __restrict__ in, float4* __restrict__ out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const int launches = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::printf("hipMalloc failed\n");
        return 1;
    }
    (void)hipMemset(a, 0, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int i = 0; i < launches; i++) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        std::printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
