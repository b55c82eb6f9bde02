// singularity_amd/csrc/post.hip — the post passes of the HDR path: bloom
// extract (2x2 Karis average + soft knee), 7-tap separable Gaussian (H then V,
// clamp-to-edge) at half resolution, and the composite (hdr + bloom*strength)
// * exposure -> ACES -> sRGB -> RGBA8.  Math follows
// render/vulkan/shaders/bloom.hlsl:8-77 and blit.hlsl:43-58; storage is
// RGBA16F between passes as in the reference's realtime targets
// (render/vulkan/vulkan_backend.cpp:1037-1050).  All passes are HBM-bound
// streaming kernels: each texel is read once per pass with 16-byte or 8-byte
// coalesced accesses, the horizontal blur stages its row segment (plus halo)
// in LDS and the vertical blur slides a 7-row register window down a column
// strip.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include "physics.hpp"
#include "post.hpp"

namespace sing {

namespace {

__constant__ float kBloomW[7] = {0.071303f, 0.131514f, 0.189879f, 0.214607f, 0.189879f, 0.131514f, 0.071303f};

__device__ __forceinline__ float3 unpack_h4(uint2 v) {
    const __half2 ab = *reinterpret_cast<const __half2*>(&v.x);
    const __half2 cd = *reinterpret_cast<const __half2*>(&v.y);
    const float2 f0 = __half22float2(ab);
    return make_float3(f0.x, f0.y, __low2float(cd));
}
__device__ __forceinline__ uint2 pack_h4(float r, float g, float b) {
    const __half2 ab = __floats2half2_rn(r, g);
    const __half2 cd = __floats2half2_rn(b, 1.0f);
    return make_uint2(*reinterpret_cast<const unsigned*>(&ab), *reinterpret_cast<const unsigned*>(&cd));
}
__device__ __forceinline__ float lum3(float3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// bloom.hlsl:19-47
__global__ __launch_bounds__(256) void extract_kernel(const uint2* __restrict__ hdr, std::uint32_t W,
                                                      std::uint32_t Wb, std::uint32_t Hb, float thr,
                                                      uint2* __restrict__ out) {
    const std::uint32_t x = blockIdx.x * 256u + threadIdx.x, y = blockIdx.y;
    if (x >= Wb || y >= Hb) return;
    const uint4 top = *reinterpret_cast<const uint4*>(hdr + std::size_t(2u * y) * W + 2u * x);
    const uint4 bot = *reinterpret_cast<const uint4*>(hdr + std::size_t(2u * y + 1u) * W + 2u * x);
    const float3 a = unpack_h4(make_uint2(top.x, top.y)), b = unpack_h4(make_uint2(top.z, top.w));
    const float3 c = unpack_h4(make_uint2(bot.x, bot.y)), d = unpack_h4(make_uint2(bot.z, bot.w));
    const float la = 1.0f / (1.0f + lum3(a)), lb = 1.0f / (1.0f + lum3(b));
    const float lc = 1.0f / (1.0f + lum3(c)), ld = 1.0f / (1.0f + lum3(d));
    const float den = fmaxf(((la + lb) + lc) + ld, 1e-6f);
    float3 avg;
    avg.x = (((a.x * la + b.x * lb) + c.x * lc) + d.x * ld) / den;
    avg.y = (((a.y * la + b.y * lb) + c.y * lc) + d.y * ld) / den;
    avg.z = (((a.z * la + b.z * lb) + c.z * lc) + d.z * ld) / den;
    const float lum = lum3(avg);
    const float knee = (lum * lum) / fmaxf(lum + thr, 1e-4f);
    const float k = (lum > 1e-4f) ? (knee / lum) : 0.0f;
    out[std::size_t(y) * Wb + x] = pack_h4(avg.x * k, avg.y * k, avg.z * k);
}

// bloom.hlsl:49-63: one block = 256 consecutive texels of one row + 3-texel halo in LDS.
__global__ __launch_bounds__(256) void blur_h_kernel(const uint2* __restrict__ in, std::uint32_t Wb,
                                                     std::uint32_t Hb, uint2* __restrict__ out) {
    __shared__ float3 tile[256 + 6];
    const std::uint32_t y = blockIdx.y, x0 = blockIdx.x * 256u;
    const uint2* row = in + std::size_t(y) * Wb;
    for (std::uint32_t i = threadIdx.x; i < 262u; i += 256u) {
        int sx = int(x0) + int(i) - 3;
        sx = sx < 0 ? 0 : (sx > int(Wb) - 1 ? int(Wb) - 1 : sx);
        tile[i] = unpack_h4(row[sx]);
    }
    __syncthreads();
    const std::uint32_t x = x0 + threadIdx.x;
    if (x >= Wb) return;
    float3 s = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const float3 v = tile[threadIdx.x + i];
        s.x = s.x + v.x * kBloomW[i];
        s.y = s.y + v.y * kBloomW[i];
        s.z = s.z + v.z * kBloomW[i];
    }
    out[std::size_t(y) * Wb + x] = pack_h4(s.x, s.y, s.z);
}

// bloom.hlsl:65-77: each thread walks a column strip of kRows outputs with a
// 7-row register window (coalesced row reads across the wave).
constexpr int kRows = 32;
__global__ __launch_bounds__(256) void blur_v_kernel(const uint2* __restrict__ in, std::uint32_t Wb,
                                                     std::uint32_t Hb, uint2* __restrict__ out) {
    const std::uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= Wb) return;
    const int y0 = int(blockIdx.y) * kRows;
    const int hmax = int(Hb) - 1;
    float3 w[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        int sy = y0 + i - 3;
        sy = sy < 0 ? 0 : (sy > hmax ? hmax : sy);
        w[i] = unpack_h4(in[std::size_t(sy) * Wb + x]);
    }
#pragma unroll
    for (int j = 0; j < kRows; ++j) {
        const int y = y0 + j;
        if (y > hmax) break;
        float3 s = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            s.x = s.x + w[i].x * kBloomW[i];
            s.y = s.y + w[i].y * kBloomW[i];
            s.z = s.z + w[i].z * kBloomW[i];
        }
        out[std::size_t(y) * Wb + x] = pack_h4(s.x, s.y, s.z);
#pragma unroll
        for (int i = 0; i < 6; ++i) w[i] = w[i + 1];
        int sy = y + 4;
        sy = sy > hmax ? hmax : sy;
        w[6] = unpack_h4(in[std::size_t(sy) * Wb + x]);
    }
}

// blit.hlsl:52-58 with the half-res bloom sampled bilinearly (clamp-to-edge)
// at the full-res pixel centre.
__global__ __launch_bounds__(256) void composite_kernel(const uint2* __restrict__ hdr, const uint2* __restrict__ bloom,
                                                        std::uint32_t W, std::uint32_t H, std::uint32_t Wb,
                                                        std::uint32_t Hb, float strength, float exposure,
                                                        uchar4* __restrict__ rgba) {
    const std::uint32_t x = blockIdx.x * 256u + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const float px = (float(x) + 0.5f) * float(Wb) / float(W) - 0.5f;
    const float py = (float(y) + 0.5f) * float(Hb) / float(H) - 0.5f;
    const float fx0 = floorf(px), fy0 = floorf(py);
    const float fx = px - fx0, fy = py - fy0;
    const int wm = int(Wb) - 1, hm = int(Hb) - 1;
    const int ix = int(fx0), iy = int(fy0);
    const int x0 = ix < 0 ? 0 : (ix > wm ? wm : ix), x1 = ix + 1 < 0 ? 0 : (ix + 1 > wm ? wm : ix + 1);
    const int y0 = iy < 0 ? 0 : (iy > hm ? hm : iy), y1 = iy + 1 < 0 ? 0 : (iy + 1 > hm ? hm : iy + 1);
    const float3 b00 = unpack_h4(bloom[std::size_t(y0) * Wb + x0]), b01 = unpack_h4(bloom[std::size_t(y0) * Wb + x1]);
    const float3 b10 = unpack_h4(bloom[std::size_t(y1) * Wb + x0]), b11 = unpack_h4(bloom[std::size_t(y1) * Wb + x1]);
    float3 t, bo, b;
    t.x = b00.x + fx * (b01.x - b00.x); t.y = b00.y + fx * (b01.y - b00.y); t.z = b00.z + fx * (b01.z - b00.z);
    bo.x = b10.x + fx * (b11.x - b10.x); bo.y = b10.y + fx * (b11.y - b10.y); bo.z = b10.z + fx * (b11.z - b10.z);
    b.x = t.x + fy * (bo.x - t.x); b.y = t.y + fy * (bo.y - t.y); b.z = t.z + fy * (bo.z - t.z);
    const float3 h = unpack_h4(hdr[std::size_t(y) * W + x]);
    const float cr = (h.x + b.x * strength) * exposure;
    const float cg = (h.y + b.y * strength) * exposure;
    const float cb = (h.z + b.z * strength) * exposure;
    rgba[std::size_t(y) * W + x] = make_uchar4(to_u8(aces1(cr)), to_u8(aces1(cg)), to_u8(aces1(cb)), 255);
}

}  // namespace

hipError_t launch_bloom(const uint2* hdr16, std::uint32_t W, std::uint32_t H, float threshold, float strength,
                        float exposure, uint2* bloom_a, uint2* bloom_b, uchar4* rgba, hipStream_t stream) {
    const std::uint32_t Wb = W / 2u, Hb = H / 2u;
    if (Wb == 0u || Hb == 0u) return hipErrorInvalidValue;
    const std::uint32_t bx = (Wb + 255u) / 256u;
    hipLaunchKernelGGL(extract_kernel, dim3(bx, Hb), dim3(256), 0, stream, hdr16, W, Wb, Hb, threshold, bloom_a);
    hipLaunchKernelGGL(blur_h_kernel, dim3(bx, Hb), dim3(256), 0, stream, bloom_a, Wb, Hb, bloom_b);
    hipLaunchKernelGGL(blur_v_kernel, dim3(bx, (Hb + kRows - 1u) / kRows), dim3(256), 0, stream, bloom_b, Wb, Hb,
                       bloom_a);
    hipLaunchKernelGGL(composite_kernel, dim3((W + 255u) / 256u, H), dim3(256), 0, stream, hdr16, bloom_a, W, H, Wb,
                       Hb, strength, exposure, rgba);
    return hipGetLastError();
}

}  // namespace sing
