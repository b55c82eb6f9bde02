// singularity_amd/csrc/batch.hpp — device-side layout of a batched render.
//
// Scenes are ordered by metric (Schwarzschild first) and split into groups
// whose termination records fit the record budget.  Within a group, scene i's
// records are [rec_base_i, rec_base_i + spp_i * P_i) (index s * P_i + pixel
// relative to rec_base_i) and its deferred (FP64) list lives at the same
// offsets of the band array.  Outputs of scene i are pixels [pix_base_i,
// pix_base_i + P_i) of the batch-wide output arenas.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "scene_const.hpp"
#include "trace.hpp"

namespace sing {

struct BatchScene {
    SceneConst<float> sf;
    SceneConst<double> sd;
    std::uint64_t rec_base;  // relative to the group's record buffer
    std::uint64_t pix_base;  // batch-wide
    std::uint64_t n_items;   // spp * n_tiles * 64 work items
    std::uint32_t P, tiles_x, n_tiles, spp;
};

// Per-scene work queues, zeroed by a fill kernel before each group.
struct SceneQueue {
    std::uint64_t next_item;   // FP32 (or FP64) trace queue
    std::uint64_t next_band;   // FP64 pass queue
    std::uint32_t band_count;  // deferred records at the front of the scene's FP64 list
    std::uint32_t band_back;   // short hard rays at its end (filled backwards; taken after the front)
};

struct BatchOutputs {
    uchar4* rgba;
    float* hdr;
    std::uint8_t* cls;
};

// One group: scenes [first, first + count) of the table, of which the first
// n_schw are Schwarzschild and the rest Kerr.
struct BatchGroup {
    std::uint32_t first, count, n_schw;
    std::uint64_t rays, pixels, pix_base;
};

struct BatchLaunch {
    const BatchScene* scenes;      // device table (all groups)
    const std::uint64_t* rec_bases;  // device, per scene (for searches)
    const std::uint64_t* pix_bases;  // device, per scene
    SceneQueue* queues;            // device, per scene
    const float2* halton;
    Rec* rec;
    std::uint32_t* band;
    QueueCtl* ctl;
};

// How waves claim rays from a scene's queue: only when at least
// rays_per_refill of their lanes are idle (or none is live), and, with
// queue_chunk > 1, from a private run of queue_chunk work items per atomic
// (<= 1: exactly one item per idle lane per atomic).
struct BatchClaim {
    std::uint32_t rays_per_refill;
    std::uint32_t queue_chunk;
};

// A batch launch's tunables (0 = not set).
struct BatchTune {
    std::uint32_t step_batch, waves_per_cu, rays_per_refill, queue_chunk;
};

// Zero the queues of scenes [first, first + count) and the ray counters
// (a kernel, so the sequence can be recorded in a graph).
hipError_t launch_batch_reset(const BatchLaunch& L, const BatchGroup& g, bool reset_totals, hipStream_t stream);
// Trace every ray of the group (mode = SING_PREC_*), FP64 pass included.
// sample_div > 1 traces only the first 1/sample_div of each scene's work
// items (autotuning).  adaptive: Kerr scenes use error-controlled steps.
hipError_t launch_batch_trace(const BatchLaunch& L, const BatchGroup& g, std::uint32_t mode, bool adaptive,
                              const LaunchShape& shape, const BatchClaim& claim, std::uint32_t sample_div,
                              bool starfield_any, hipStream_t stream);
hipError_t launch_batch_shade(const BatchLaunch& L, const BatchGroup& g, const BatchOutputs& out,
                              hipStream_t stream);

}  // namespace sing
