// singularity_amd/csrc/trace.hpp — host-visible declarations of the frame pipeline.
//
// A frame is rendered in sample chunks.  For each chunk of S samples:
//   1. trace   (FP32 or FP64): persistent wave64 kernel, one ray per lane with
//      lane refill; writes one 16-byte termination record per (sample, pixel).
//      In mixed precision the FP32 trace defers photon-band rays to a list.
//   2. trace64 over the deferred list (mixed precision only).
//   3. shade: one thread per pixel reads its S records in sample order, shades
//      them, tone-maps each sample and accumulates; after the last chunk it
//      writes the frame outputs.
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

// One termination record per traced sample.
//   escape: a = theta, b = phi            disc: a = r_cross, b = phi, c = L_z (Kerr) or u^phi (Schwarzschild)
//   meta  = term (bits 0..2) | deferred (bit 3) | steps << 4
struct Rec {
    float a, b, c;
    std::uint32_t meta;
};
constexpr std::uint32_t kRecDeferred = 8u;

// Running per-pixel sums between chunks.
struct PixAcc {
    float ldr[3];
    float hdr[3];
    std::uint32_t steps;
    std::uint32_t cls;
};

// Device-side work queues and counters, zeroed before each chunk.
struct QueueCtl {
    std::uint64_t next_item;   // trace queue
    std::uint64_t next_band;   // FP64 pass queue
    std::uint32_t band_count;  // deferred records at the front of the FP64 list (long hard rays, FP32-pass deferrals)
    std::uint32_t band_back;   // short hard rays, at the back of the list (filled from its end)
    std::uint64_t steps;
    std::uint64_t rays;
    std::uint64_t rays64;
    std::uint64_t rejected;    // error-controlled stepping: rejected step attempts
    std::uint64_t steps32;     // RK4 steps executed by the FP32 trace pass (deferred rays included)
    std::uint64_t steps64;     // RK4 steps executed by the FP64 trace pass
};

// Work item of an FP32 pass that needs no trace (tile padding outside the
// image, or a ray whose record the set-up kernel already wrote).
constexpr std::uint32_t kItemSkip = 0xFFFFFFFFu;

struct Chunk {
    Rec* rec;                   // S * P records, index s_local * P + pixel
    std::uint32_t* band;        // deferred record indices (capacity S * P)
    // FP32 passes, per work item (S * n_tiles * 64): the ray's initial
    // (v0, v1, v2, v3) of RayState — Kerr (p_r, p_theta, E, L_z),
    // Schwarzschild (u^t, u^r, u^theta, u^phi) — and its record index, or
    // kItemSkip; written by the set-up kernel before the trace
    float4* init;
    std::uint32_t* item_rec;
    std::uint32_t s0, S;        // first sample and sample count of the chunk
    std::uint32_t P;            // pixels
    std::uint32_t tiles_x, n_tiles;
    // the FP64 passes' scene constants in device memory: a ray's set-up reads
    // them where it runs instead of holding ~90 scalar registers of kernel
    // arguments live across the persistent loop
    const SceneConst<double>* sd;
};

struct LaunchShape {
    std::uint32_t blocks;      // persistent grid, 256 threads per block
    std::uint32_t blocks64;    // persistent grid of the FP64 kernels
    std::uint32_t step_batch;  // RK4 steps between refills
};

// mode: SING_PREC_FP32 -> FP32 trace; SING_PREC_FP64 -> FP64 trace of every
// ray; SING_PREC_MIXED -> FP32 trace with band deferral, then FP64 over the list.
// ev (optional, 4 events): recorded before and after the FP32 trace kernel
// (ev[0], ev[1]) and before and after the FP64 trace kernel (ev[2], ev[3]), on
// the stream that runs it; a pass the mode does not run gets its two events
// back to back.  tail (optional, mixed only): the stream of the FP64 list
// pass; `join` is recorded on `stream` after the FP32 pass and its audit, and
// `tail` waits on it, so the list pass can run beside whatever `stream` is
// given next.
hipError_t launch_trace_chunk(const SceneConst<float>& sf, const SceneConst<double>& sd, const float2* halton,
                              const Chunk& ch, QueueCtl* ctl, std::uint32_t mode, const LaunchShape& shape,
                              hipStream_t stream, const hipEvent_t* ev = nullptr, hipStream_t tail = nullptr,
                              hipEvent_t join = nullptr);
hipError_t launch_shade_chunk(const SceneConst<float>& sf, const Chunk& ch, PixAcc* acc, bool first, bool last,
                              const TraceOutputs& out, hipStream_t stream);

}  // namespace sing
