// singularity_amd/csrc/hbm_copy.hip — HBM counter calibration: copies 256 MiB
// three times with a grid-stride 16-byte kernel, so each dispatch moves exactly
// 256 MiB in and 256 MiB out.  Profiling it with rocprofv3 --pmc FETCH_SIZE
// (or WRITE_SIZE) gives the factor between the counter and the bytes moved;
// the same factor converts the trace kernels' counters into HBM bytes.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>

__global__ __launch_bounds__(256) void hbm_copy_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                       std::size_t n) {
    for (std::size_t i = std::size_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += std::size_t(gridDim.x) * 256u)
        b[i] = a[i];
}

int main() {
    const std::size_t bytes = std::size_t(256) << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "hbm_copy: hipMalloc failed\n");
        return 1;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(1024), dim3(256), 0, 0, a, b, n);  // first touch
    (void)hipEventRecord(e0, 0);
    for (int it = 0; it < 3; ++it) hipLaunchKernelGGL(hbm_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "hbm_copy: kernel failed\n");
        return 2;
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::printf("hbm_copy: %zu bytes read + %zu written per dispatch, %.1f GB/s\n", bytes, bytes,
                3.0 * 2.0 * double(bytes) / (ms * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
