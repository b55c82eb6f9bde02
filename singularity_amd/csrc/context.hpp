// singularity_amd/csrc/context.hpp — the render context behind the C ABI's
// opaque SingContext (shared by the frame, adjoint and batch entry points).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sing_abi.h"
#include "adjoint.hpp"
#include "batch.hpp"
#include "trace.hpp"

void sing_set_global_error(const std::string& m);

namespace sing {
// Checks the options and uniform block of a frame render; on failure writes
// "<who>: <reason>" into msg and returns false.
bool render_args_ok(const SingUniforms& u, const SingRenderOpts& o, const char* who, char* msg, std::size_t len);
// Persistent grid and step batch of a frame's trace kernels on a device with num_cu CUs.
LaunchShape launch_shape_for(int num_cu, const SingRenderOpts& o);
}  // namespace sing

struct SingContext {
    int device = 0;
    int num_cu = 256;
    std::uint32_t width = 0, height = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // second stream (default priority) for concurrent passes
    std::vector<hipEvent_t> events;       // timing events
    std::vector<hipEvent_t> join_events;  // cross-stream joins (created without timing)
    // frame outputs
    uchar4* d_rgba = nullptr;
    float* d_hdr = nullptr;
    uint2* d_hdr16 = nullptr;
    std::uint8_t* d_cls = nullptr;
    std::uint32_t* d_steps = nullptr;
    sing::PixAcc* d_acc = nullptr;
    // per-chunk ray records and the deferred (FP64) list
    sing::Rec* d_rec = nullptr;
    std::uint32_t* d_band = nullptr;
    std::size_t rec_capacity = 0;
    // per work item of an FP32 pass: the set-up kernel's ray init and record index
    float4* d_init = nullptr;
    std::uint32_t* d_item_rec = nullptr;
    std::size_t item_capacity = 0;
    sing::QueueCtl* d_ctl = nullptr;
    float2* d_halton = nullptr;
    sing::SceneConst<double>* d_sd = nullptr;  // device copy of the frame's FP64 scene constants
    sing::SceneConst<double> h_sd{};           // its host source (lives until the copy has run)
    std::uint32_t outputs_valid = 0;  // SING_OUT_* of the last frame
    std::string err;
    // a frame submitted by sing_render_async and not yet completed: what its
    // completion needs (stats from the events and the queue counters)
    struct PendingFrame {
        std::uint32_t n_chunks, nbuf, outs, precision, post;
        bool overlap;
        std::uint64_t expect_rays;
    };
    bool pending = false;
    PendingFrame pend{};
    bool stats_unclaimed = false;      // a completed async frame whose stats sing_wait has not returned
    SingRenderStats last_stats{};
    int last_rc = 0;
    sing::QueueCtl* h_ctl = nullptr;   // pinned host copy of the two queue-counter blocks
    // differentiable render
    sing::DiffSet* d_dset = nullptr;  // one per evaluation of a call
    sing::DiffBuffers diff{};
    float* d_target = nullptr;
    float* d_weight = nullptr;
    double* d_out = nullptr;  // the adjoint's reduced results (sing::kDiffOut doubles per evaluation)
    double* h_out = nullptr;  // pinned host copy of d_out
    std::size_t diff_P = 0, diff_n = 0, diff_ncp = 0, diff_C = 0;  // diff_P: pixels of one image
    // batched render
    sing::BatchScene* d_bscenes = nullptr;
    std::uint64_t* d_brec_bases = nullptr;
    std::uint64_t* d_bpix_bases = nullptr;
    sing::SceneQueue* d_bqueues = nullptr;
    std::size_t b_scene_cap = 0, b_rb_cap = 0, b_pb_cap = 0, b_q_cap = 0;
    sing::Rec* d_brec = nullptr;
    std::uint32_t* d_bband = nullptr;
    std::size_t b_rec_cap = 0, b_band_cap = 0;
    uchar4* d_brgba = nullptr;
    float* d_bhdr = nullptr;
    std::uint8_t* d_bcls = nullptr;
    std::size_t b_rgba_cap = 0, b_hdr_cap = 0, b_cls_cap = 0;
    hipGraph_t b_graph = nullptr;
    hipGraphExec_t b_graph_exec = nullptr;
    std::vector<std::uint64_t> b_graph_key;
    std::vector<std::pair<std::uint64_t, sing::BatchTune>> b_tuned;  // kind of batch -> tuned launch

    int fail(int code, const char* fmt, ...);
    // Completes a submitted frame (waits for it, fills last_stats, checks its
    // ray total); every entry point on the context calls it first, so a
    // pending frame is never overwritten or read half-done.  Returns the
    // frame's status (SING_OK when nothing was pending).
    int complete_pending();
    int hip_fail(hipError_t e, const char* where);
    void free_frame();
    int alloc_frame(std::uint32_t w, std::uint32_t h);
    int ensure_records(std::size_t rays);
    int ensure_items(std::size_t items);
    int ensure_events(std::size_t n);
    int ensure_join_events(std::size_t n);
    std::uint32_t chunk_samples(std::uint32_t spp) const;
    sing::LaunchShape launch_shape(const SingRenderOpts& o) const;
    int ensure_diff(std::size_t Pimg, std::size_t n_eval, std::size_t ncp, std::size_t C);
    void free_diff();
    void free_batch();
    void drop_batch_graph();
};
