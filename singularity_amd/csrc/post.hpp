// singularity_amd/csrc/post.hpp — bloom + composite post passes (host launch).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sing {

// hdr16: W*H RGBA16F (16-byte aligned); the half-res bloom texels live in LDS.
// rgba: W*H RGBA8 composite (hdr + bloom*strength)*exposure -> ACES -> sRGB.
hipError_t launch_bloom(const uint2* hdr16, std::uint32_t W, std::uint32_t H, float threshold, float strength,
                        float exposure, uchar4* rgba, hipStream_t stream);

}  // namespace sing
