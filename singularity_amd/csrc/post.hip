// singularity_amd/csrc/post.hip — the post passes of the HDR path: bloom
// extract (2x2 Karis average + soft knee), 7-tap separable Gaussian (H then V,
// clamp-to-edge) at half resolution, and the composite (hdr + bloom*strength)
// * exposure -> ACES -> sRGB -> RGBA8.  Math follows
// render/vulkan/shaders/bloom.hlsl:8-77 and blit.hlsl:43-58; storage is
// RGBA16F between passes as in the reference's realtime targets
// (render/vulkan/vulkan_backend.cpp:1037-1050), here kept in LDS: the four
// passes run as one tiled kernel that reads the HDR input and writes the RGBA8
// output and nothing else (the post pass is HBM-bound; the reference's four
// full-image passes move 32 B per full-res pixel, this one about 19).
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

// bloom.hlsl:19-47 at half-res texel (col, row): 2x2 Karis average of the
// full-res HDR block and the soft knee, rounded to RGBA16F storage.
__device__ __forceinline__ uint2 extract_texel(const uint2* __restrict__ hdr, std::uint32_t W, std::uint32_t col,
                                               std::uint32_t row, float thr) {
    const uint4 top = *reinterpret_cast<const uint4*>(hdr + std::size_t(2u * row) * W + 2u * col);
    const uint4 bot = *reinterpret_cast<const uint4*>(hdr + std::size_t(2u * row + 1u) * W + 2u * col);
    const float3 a = unpack_h4(make_uint2(top.x, top.y)), b = unpack_h4(make_uint2(top.z, top.w));
    const float3 c = unpack_h4(make_uint2(bot.x, bot.y)), d = unpack_h4(make_uint2(bot.z, bot.w));
    // divisions as hardware reciprocals (a few float ulps, far below the half-float
    // rounding of the stored result)
    const float la = __builtin_amdgcn_rcpf(1.0f + lum3(a)), lb = __builtin_amdgcn_rcpf(1.0f + lum3(b));
    const float lc = __builtin_amdgcn_rcpf(1.0f + lum3(c)), ld = __builtin_amdgcn_rcpf(1.0f + lum3(d));
    const float inv = __builtin_amdgcn_rcpf(fmaxf(((la + lb) + lc) + ld, 1e-6f));
    float3 avg;
    avg.x = (((a.x * la + b.x * lb) + c.x * lc) + d.x * ld) * inv;
    avg.y = (((a.y * la + b.y * lb) + c.y * lc) + d.y * ld) * inv;
    avg.z = (((a.z * la + b.z * lb) + c.z * lc) + d.z * ld) * inv;
    const float lum = lum3(avg);
    // knee / lum = lum / (lum + thr)
    const float k = (lum > 1e-4f) ? lum * __builtin_amdgcn_rcpf(fmaxf(lum + thr, 1e-4f)) : 0.0f;
    return pack_h4(avg.x * k, avg.y * k, avg.z * k);
}

// RGBA16F storage of a pass's output: round to half and back.
__device__ __forceinline__ float3 store_h(float r, float g, float b) { return unpack_h4(pack_h4(r, g, b)); }

// 7-tap blur (bloom.hlsl:49-77) of seven texels `stride` apart, rounded to RGBA16F.
__device__ __forceinline__ float3 blur7(const float3* v, int stride) {
    float3 s = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const float3 t = v[i * stride];
        s.x = s.x + t.x * kBloomW[i];
        s.y = s.y + t.y * kBloomW[i];
        s.z = s.z + t.z * kBloomW[i];
    }
    return store_h(s.x, s.y, s.z);
}

// Composite sample position of full-res pixel x (blit.hlsl's bilinear fetch
// of the half-res bloom at the pixel centre): floor index and fraction.
__device__ __forceinline__ void bloom_coord(std::uint32_t x, std::uint32_t W, std::uint32_t Wb, int& i0, float& f) {
    const float p = (float(x) + 0.5f) * float(Wb) / float(W) - 0.5f;
    const float fl = floorf(p);
    f = p - fl;
    i0 = int(fl);
}

// The composite's tone map with the hardware's transcendental instructions:
// Narkowicz ACES with one reciprocal, the sRGB curve's power as exp2(log2(c) /
// 2.4), then 8 bits.  Each is within a few float ulps of the exact form (the
// shade kernel's aces1/to_u8), far below the half-LSB of an 8-bit channel.
__device__ __forceinline__ std::uint8_t tone_u8(float x) {
    float a = (x * (2.51f * x + 0.03f)) * __builtin_amdgcn_rcpf(x * (2.43f * x + 0.59f) + 0.14f);
    a = fminf(fmaxf(a, 0.0f), 1.0f);
    const float s = a <= 0.0031308f ? 12.92f * a
                                    : 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(a) * (1.0f / 2.4f)) - 0.055f;
    return std::uint8_t(fminf(fmaxf(s * 255.0f + 0.5f, 0.0f), 255.0f));
}

// The whole bloom post pass in one launch.  One block renders a kTW x kTH tile
// of the RGBA8 output: it extracts the half-res texels the tile's bilinear
// fetches and the two 7-tap passes reach (the tile's half-res footprint plus
// one texel for the bilinear fetch and three for each blur, clamped to the
// image), blurs them horizontally and vertically in LDS, and composites.  The
// intermediates are rounded to RGBA16F exactly where the reference's passes
// store them (and kept in LDS as floats, so each is converted once); nothing
// but the HDR input and the RGBA8 output touches HBM (halo texels are
// extracted by each tile that needs them).  The bilinear positions, which
// cost a division each, are computed once per tile column and row.  Tiles
// are dealt to the 8 XCDs in contiguous runs (block b runs on XCD b % 8), so
// neighbouring tiles, which read the same halo rows, share an L2.
constexpr int kTW = 64, kTH = 32;             // output tile (full-res pixels)
constexpr int kVW = kTW / 2 + 2, kVH = kTH / 2 + 2;  // vertical-blur texels a tile's bilinear fetches read
constexpr int kHW = kVW, kHH = kVH + 6;       // horizontal-blur texels the vertical pass reads
constexpr int kEW = kHW + 6, kEH = kHH;       // extracted texels the horizontal pass reads
constexpr int kXcd = 8;
constexpr int kThreads = 512;

__global__ __launch_bounds__(kThreads) void bloom_kernel(const uint2* __restrict__ hdr, std::uint32_t W, std::uint32_t H,
                                                    std::uint32_t Wb, std::uint32_t Hb, float thr, float strength,
                                                    float exposure, std::uint32_t tiles_x, std::uint32_t n_tiles,
                                                    uchar4* __restrict__ rgba) {
    __shared__ float3 ext[kEH][kEW];
    __shared__ float3 bh[kHH][kHW];
    __shared__ float3 bv[kVH][kVW];
    __shared__ int col0[kTW], col1[kTW], row0[kTH], row1[kTH];
    __shared__ float colf[kTW], rowf[kTH];
    // tile order: the n_tiles tiles in kXcd contiguous runs, run k to the blocks b with b % kXcd == k
    const std::uint32_t b = blockIdx.x, q = n_tiles / kXcd, r = n_tiles % kXcd;
    const std::uint32_t xcd = b % kXcd, idx = b / kXcd;
    const std::uint32_t tile = xcd < r ? xcd * (q + 1u) + idx : r * (q + 1u) + (xcd - r) * q + idx;
    const std::uint32_t X0 = (tile % tiles_x) * kTW, Y0 = (tile / tiles_x) * kTH;
    const int wm = int(Wb) - 1, hm = int(Hb) - 1;
    int cx, cy;
    float unused;
    bloom_coord(X0, W, Wb, cx, unused);
    bloom_coord(Y0, H, Hb, cy, unused);
    const int hx = cx < 0 ? 0 : cx, hy = cy < 0 ? 0 : cy;  // first vertical-blur texel of the tile
    const int t = int(threadIdx.x);
    if (t < kTW) {
        int ix;
        float fx;
        bloom_coord(X0 + std::uint32_t(t), W, Wb, ix, fx);
        col0[t] = min(max(ix, 0), wm) - hx;
        col1[t] = min(max(ix + 1, 0), wm) - hx;
        colf[t] = fx;
    } else if (t < kTW + kTH) {
        int iy;
        float fy;
        bloom_coord(Y0 + std::uint32_t(t - kTW), H, Hb, iy, fy);
        row0[t - kTW] = min(max(iy, 0), hm) - hy;
        row1[t - kTW] = min(max(iy + 1, 0), hm) - hy;
        rowf[t - kTW] = fy;
    }
    for (int i = t; i < kEH * kEW; i += kThreads) {
        const int er = i / kEW, ec = i - er * kEW;
        const int row = min(max(hy - 3 + er, 0), hm), col = min(max(hx - 3 + ec, 0), wm);
        ext[er][ec] = unpack_h4(extract_texel(hdr, W, std::uint32_t(col), std::uint32_t(row), thr));
    }
    __syncthreads();
    for (int i = t; i < kHH * kHW; i += kThreads) {
        const int rr = i / kHW, cc = i - rr * kHW;
        bh[rr][cc] = blur7(&ext[rr][cc], 1);
    }
    __syncthreads();
    for (int i = t; i < kVH * kVW; i += kThreads) {
        const int rr = i / kVW, cc = i - rr * kVW;
        bv[rr][cc] = blur7(&bh[rr][cc], kHW);
    }
    __syncthreads();
    for (int i = t; i < kTW * kTH; i += kThreads) {
        const int lx = i % kTW, ly = i / kTW;
        const std::uint32_t x = X0 + std::uint32_t(lx), y = Y0 + std::uint32_t(ly);
        if (x >= W || y >= H) continue;
        const int x0 = col0[lx], x1 = col1[lx], y0 = row0[ly], y1 = row1[ly];
        const float fx = colf[lx], fy = rowf[ly];
        const float3 b00 = bv[y0][x0], b01 = bv[y0][x1], b10 = bv[y1][x0], b11 = bv[y1][x1];
        float3 tp, bo, bl;
        tp.x = b00.x + fx * (b01.x - b00.x); tp.y = b00.y + fx * (b01.y - b00.y); tp.z = b00.z + fx * (b01.z - b00.z);
        bo.x = b10.x + fx * (b11.x - b10.x); bo.y = b10.y + fx * (b11.y - b10.y); bo.z = b10.z + fx * (b11.z - b10.z);
        bl.x = tp.x + fy * (bo.x - tp.x); bl.y = tp.y + fy * (bo.y - tp.y); bl.z = tp.z + fy * (bo.z - tp.z);
        const float3 h = unpack_h4(hdr[std::size_t(y) * W + x]);
        const float cr = (h.x + bl.x * strength) * exposure;
        const float cg = (h.y + bl.y * strength) * exposure;
        const float cb = (h.z + bl.z * strength) * exposure;
        rgba[std::size_t(y) * W + x] = make_uchar4(tone_u8(cr), tone_u8(cg), tone_u8(cb), 255);
    }
}

}  // namespace

hipError_t launch_bloom(const uint2* hdr16, std::uint32_t W, std::uint32_t H, float threshold, float strength,
                        float exposure, uchar4* rgba, hipStream_t stream) {
    const std::uint32_t Wb = W / 2u, Hb = H / 2u;
    if (Wb == 0u || Hb == 0u) return hipErrorInvalidValue;
    const std::uint32_t tiles_x = (W + kTW - 1u) / kTW, tiles_y = (H + kTH - 1u) / kTH;
    hipLaunchKernelGGL(bloom_kernel, dim3(tiles_x * tiles_y), dim3(kThreads), 0, stream, hdr16, W, H, Wb, Hb, threshold,
                       strength, exposure, tiles_x, tiles_x * tiles_y, rgba);
    return hipGetLastError();
}

}  // namespace sing
