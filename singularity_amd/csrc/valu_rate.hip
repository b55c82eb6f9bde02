// singularity_amd/csrc/valu_rate.hip — VALU issue-rate calibration for the
// trace kernels' roofline.  Each kernel runs 8 independent dependency chains
// per lane of one instruction class (FP32 FMA and multiply with scalar,
// vector-register or literal operands, FP32 reciprocal, FP64 FMA) or alternating between
// two (multiply and FMA), over a grid of 16 four-wave blocks per CU, with no
// memory traffic in the loop, and reports lane-instructions per second and per
// CU per shader clock.  The shader clock is read in block 0 as the ratio of
// clock64() ticks to the 100 MHz constant wall clock over the kernel.  A
// result of 64 lanes per CU-clock means one wave64 instruction per CU per
// clock.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kChains = 8;
constexpr int kIters = 4096;

__device__ unsigned long long g_ticks[2];

// A register copy the optimiser cannot see through, so that per-lane operands
// stay in vector registers and are not folded into constants.
__device__ __forceinline__ float opaque(float x) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ double opaque(double x) { return x; }

template <class T, int OP>
__global__ __launch_bounds__(256) void rate_kernel(T* __restrict__ out, T a, T b) {
    unsigned long long c0 = 0, w0 = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c0 = clock64();
        w0 = wall_clock64();
    }
    T x[kChains], yv[kChains], zv[kChains];
    for (int c = 0; c < kChains; ++c) {
        x[c] = a + T(threadIdx.x + c);
        yv[c] = opaque(a - T(1e-4) * T((threadIdx.x + c) & 7));  // per-lane operands held in VGPRs
        zv[c] = opaque(b * T(1 + ((threadIdx.x + c) & 3)));
    }
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if constexpr (OP == 0) x[c] = x[c] * a + b;  // FMA
            if constexpr (OP == 1) x[c] = x[c] * a;      // multiply
            if constexpr (OP == 2) x[c] = __builtin_amdgcn_rcpf(x[c]);
            if constexpr (OP == 4) x[c] = x[c] * zv[c] + x[c];  // FMA, VGPR operands (v_fmac form)
            if constexpr (OP == 5) x[c] = x[c] * yv[c];          // multiply, VGPR operands
            if constexpr (OP == 6) x[c] = x[c] * yv[c] + T(0.1234567);  // FMA, VGPR operands and a literal
            if constexpr (OP == 7) x[c] = x[c] * yv[c] + zv[c];  // FMA, three VGPR operands (VOP3)
            if constexpr (OP == 3) {  // independent multiply and FMA chains, alternating
                if (c & 1) x[c] = x[c] * a;
                else x[c] = x[c] * b + a;
            }
        }
    }
    T s = T(0);
    for (int c = 0; c < kChains; ++c) s += x[c];
    out[std::size_t(blockIdx.x) * 256u + threadIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g_ticks[0] = clock64() - c0;
        g_ticks[1] = wall_clock64() - w0;
    }
}

template <class T, int OP>
bool run(void* buf, unsigned blocks, int cus, int wall_khz, const char* name) {
    T* out = static_cast<T*>(buf);
    hipLaunchKernelGGL((rate_kernel<T, OP>), dim3(blocks), dim3(256), 0, 0, out, T(0.999), T(0.001));  // warm-up
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return false;
    const int reps = 3;
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL((rate_kernel<T, OP>), dim3(blocks), dim3(256), 0, 0, out, T(0.999), T(0.001));
    (void)hipEventRecord(e1, 0);
    if (hipDeviceSynchronize() != hipSuccess) return false;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    unsigned long long t[2] = {0, 0};
    if (hipMemcpyFromSymbol(t, HIP_SYMBOL(g_ticks), sizeof(t)) != hipSuccess) return false;
    const double ghz = double(t[0]) / (double(t[1]) / (double(wall_khz) * 1e3)) * 1e-9;
    const double lanes = double(blocks) * 256.0 * kChains * kIters * reps;
    const double rate = lanes / (double(ms) * 1e-3);
    std::printf("%-12s %8.3f ms per launch  %7.2f T lane-instr/s  shader clock %.3f GHz  %6.1f lanes/CU/clock\n",
                name, ms / reps, rate * 1e-12, ghz, rate / (double(cus) * ghz * 1e9));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return true;
}

}  // namespace

int main() {
    int cus = 0, wall_khz = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
        hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || cus <= 0 ||
        wall_khz <= 0) {
        std::fprintf(stderr, "valu_rate: no device\n");
        return 1;
    }
    const unsigned blocks = unsigned(cus) * 16u;
    void* buf = nullptr;
    if (hipMalloc(&buf, std::size_t(blocks) * 256u * sizeof(double)) != hipSuccess) {
        std::fprintf(stderr, "valu_rate: hipMalloc failed\n");
        return 1;
    }
    std::printf("valu_rate: %d CUs, %u blocks of 256, %d chains x %d iterations per lane\n", cus, blocks, kChains,
                kIters);
    const bool ok = run<float, 0>(buf, blocks, cus, wall_khz, "fp32 fma") &&
                    run<float, 1>(buf, blocks, cus, wall_khz, "fp32 mul") &&
                    run<float, 2>(buf, blocks, cus, wall_khz, "fp32 rcp") &&
                    run<float, 3>(buf, blocks, cus, wall_khz, "mul|fma mix") &&
                    run<float, 4>(buf, blocks, cus, wall_khz, "fp32 fmac v") &&
                    run<float, 5>(buf, blocks, cus, wall_khz, "fp32 mul v") &&
                    run<float, 6>(buf, blocks, cus, wall_khz, "fp32 fma vk") &&
                    run<float, 7>(buf, blocks, cus, wall_khz, "fp32 fma vvv") &&
                    run<double, 0>(buf, blocks, cus, wall_khz, "fp64 fma");
    (void)hipFree(buf);
    if (!ok) {
        std::fprintf(stderr, "valu_rate: kernel failed\n");
        return 2;
    }
    return 0;
}
