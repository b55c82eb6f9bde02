// singularity_amd/csrc/trace.hpp — host-visible declarations of the trace launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "scene_const.hpp"

namespace sing {

// Device output pointers of one frame; any may be null.
struct TraceOutputs {
    uchar4* rgba;          // W*H RGBA8, reference semantics
    float* hdr;            // W*H*3 float, mean radiance before exposure
    uint2* hdr16;          // W*H half4 (RGBA16F), input of the post passes
    std::uint8_t* cls;     // W*H termination class of sample 0
    std::uint32_t* steps;  // W*H RK4 steps summed over samples
};

// Device-side work queue and counters, zeroed before each launch.
struct QueueCtl {
    std::uint64_t next_item;
    std::uint64_t steps;
    std::uint64_t rays;
    std::uint64_t rays64;
};

struct LaunchShape {
    std::uint32_t precision;   // SING_PREC_*
    std::uint32_t blocks;      // persistent grid, 256 threads per block
    std::uint32_t step_batch;  // RK4 steps between refills
};

hipError_t launch_trace(const SceneConst<float>& sf, const SceneConst<double>& sd, const float2* halton,
                        const TraceOutputs& out, QueueCtl* ctl, const LaunchShape& shape, hipStream_t stream);

}  // namespace sing
