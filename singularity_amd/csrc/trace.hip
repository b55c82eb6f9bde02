// singularity_amd/csrc/trace.hip — trace and shade kernels of the frame pipeline.
//
// Trace: a persistent wave64 kernel.  Each lane integrates one ray (one
// (pixel, sample) pair) at a time; after every `step_batch` RK4 steps the lanes
// whose ray terminated write its 16-byte record and the wave claims new rays
// for all of them with a single atomic, so long rays never hold the other 63
// lanes of the wave hostage.  The lean step loop carries only the ray state;
// all shading lives in the shade kernel.
//
// Shade: one thread per pixel walks its records in sample order and
// accumulates exactly like the reference's per-pixel sample loop
// (render/cuda/kernels/geodesic_kernel.cu:462-510): per-sample ACES of the
// exposed radiance, mean over samples, sRGB, RGBA8.  The fixed order keeps
// frames bit-reproducible with no float atomics.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "physics.hpp"
#include "trace.hpp"
#include "trace_core.hpp"

namespace sing {

namespace {

using namespace core;

// METRIC 0: Schwarzschild; 1: Kerr; 2: Kerr with error-controlled steps; 3: Kerr
// with the fixed step and SING_REFINE_* options (FP64 crossing root, null-
// constraint re-projection).
// PASS 0: FP32 trace of every ray; 1: FP32 trace deferring photon-band rays;
// 2: FP64 trace of every ray; 3: FP64 trace of the deferred list.
// The fixed-step Kerr FP64 passes are held to 168 VGPRs, 3 waves per SIMD (the
// compiler's own choice, 170, leaves room for 2), without spills; METRIC 3
// keeps the compiler's choice (held to 168 it spills).  The error-controlled
// FP32 passes are held to 128 VGPRs, 4 waves per SIMD (left alone the mixed
// form's pass takes 133, 3 waves; held, it spills one 16-byte value, outside
// the step loop): that pass 6.6 % faster.  The error-controlled FP64 passes
// (220 VGPRs, 2 waves) were slower held to 168 (3 waves; 176 bytes spilled,
// reloaded in the step loop).
template <int METRIC, int PASS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((METRIC == 1 && PASS >= 2) ? 3
                                                                     : (METRIC == 2 && PASS < 2) ? 4
                                                                                                 : 1)))
void trace_kernel(SceneConst<float> sf, SceneConst<double> sd,
                                                    const float2* __restrict__ halton, Chunk ch,
                                                    QueueCtl* __restrict__ ctl, std::uint32_t step_batch) {
    constexpr bool KERR = (METRIC >= 1);
    constexpr bool ADAPT = (METRIC == 2);
    constexpr bool XTRA = (METRIC == 3);
    constexpr bool F64 = (PASS >= 2);
    using S = typename std::conditional<F64, double, float>::type;
    const int lane = __lane_id();
    const std::uint32_t W = sf.width, H = sf.height, spp = sf.spp;
    // the FP64 list: band_count entries at the front, band_back at the end of
    // the chunk's S * P slots (filled backwards), taken front first
    const std::uint64_t n_front = (PASS == 3) ? std::uint64_t(ctl->band_count) : 0u;
    const std::uint64_t n_items =
        (PASS == 3) ? n_front + ctl->band_back : std::uint64_t(ch.S) * ch.n_tiles * 64u;
    std::uint64_t* counter = (PASS == 3) ? &ctl->next_band : &ctl->next_item;

    bool retired = false, live = false, need = true;
    std::uint32_t rec_idx = 0;
    std::uint64_t lane_steps = 0, lane_rays = 0, lane_exec = 0;
    std::uint32_t lane_rej = 0;
    RayState<S> y{};

    for (;;) {
        if constexpr (!F64) {
            // (1) claim set-up rays for idle lanes with one atomic per wave and
            // round; an item the set-up kernel resolved (kItemSkip) sends its
            // lane round again
            for (int round = 0; round < 4; ++round) {
                const std::uint64_t want = __ballot(need && !retired);
                if (!want) break;
                const int leader = __ffsll((unsigned long long)want) - 1;
                std::uint64_t base = 0;
                if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)__popcll(want));
                base = shfl_u64(base, leader);
                if (need && !retired) {
                    const std::uint64_t item = base + __popcll(want & ((1ull << lane) - 1ull));
                    if (item >= n_items) {
                        retired = true;
                    } else {
                        const std::uint32_t ri = ch.item_rec[item];
                        if (ri != kItemSkip) {
                            rec_idx = ri;
                            start_ray<S, KERR>(y, sf, ch.init[item]);
                            need = false;
                            live = true;
                        }
                    }
                }
            }
        } else {
            // (1) FP64 passes: claim rays for idle lanes with one atomic per
            // wave and set them up here, in double
            const std::uint64_t want = __ballot(need && !retired);
            if (want) {
                const int leader = __ffsll((unsigned long long)want) - 1;
                std::uint64_t base = 0;
                if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)__popcll(want));
                base = shfl_u64(base, leader);
                if (need && !retired) {
                    const std::uint64_t item = base + __popcll(want & ((1ull << lane) - 1ull));
                    if (item >= n_items) {
                        retired = true;
                    } else {
                        std::uint32_t gx, gy, s;
                        bool ok;
                        if (PASS == 3) {
                            rec_idx = ch.band[item < n_front ? item : std::uint64_t(ch.S) * ch.P - 1u - (item - n_front)];
                            const std::uint32_t sl = rec_idx / ch.P, pix = rec_idx - sl * ch.P;
                            gx = pix % W;
                            gy = pix / W;
                            s = ch.s0 + sl;
                            ok = true;
                        } else {
                            ok = item_to_ray(item, ch, W, H, rec_idx, gx, gy, s);
                        }
                        if (ok) {
                            float ox = 0.5f, oy = 0.5f;
                            if (spp > 1u) {
                                const float2 hv = halton[s];
                                ox = hv.x;
                                oy = hv.y;
                            }
                            const bool started = init_ray<double, S, KERR>(y, *ch.sd, gx, gy, double(ox),
                                                                           double(oy), (bool*)nullptr);
                            need = false;
                            if (!started) {
                                Rec rc;
                                rc.a = rc.b = rc.c = 0.0f;
                                rc.meta = kTermInitFail;
                                ch.rec[rec_idx] = rc;
                                lane_rays += 1;
                                need = true;
                            } else {
                                live = true;
                            }
                        }
                    }
                }
            }
        }
        if (__ballot(!retired) == 0ull) break;

        // (2) integrate a batch of steps
        std::uint32_t term;
        if constexpr (F64) term = run_steps<double, S, KERR, false, ADAPT, XTRA>(y, sd, live, step_batch, &lane_rej);
        else term = run_steps<float, S, KERR, PASS == 1, ADAPT, XTRA>(y, sf, live, step_batch, &lane_rej);

        // (3) emit finished rays; in the mixed pass, hand ambiguous outcomes to FP64
        if (term != kTermLive) {
            lane_exec += y.steps;
            bool defer = false;
            if (PASS == 1) defer = needs_fp64(y, term);
            if (defer) {
                const std::uint32_t slot = atomicAdd(&ctl->band_count, 1u);
                ch.band[slot] = rec_idx;
                ch.rec[rec_idx] = deferred_marker();
            } else {
                ch.rec[rec_idx] = make_rec<S, KERR, ADAPT, ADAPT || XTRA>(y, term);
                lane_steps += y.steps;
                lane_rays += 1;
            }
            need = true;
        }
    }

    std::uint64_t rej = lane_rej;
    for (int off = 32; off > 0; off >>= 1) {
        lane_steps += shfl_u64(lane_steps, lane ^ off);
        lane_rays += shfl_u64(lane_rays, lane ^ off);
        lane_exec += shfl_u64(lane_exec, lane ^ off);
        if (ADAPT) rej += shfl_u64(rej, lane ^ off);
    }
    if (ADAPT && lane == 0 && rej) atomicAdd((unsigned long long*)&ctl->rejected, (unsigned long long)rej);
    if (lane == 0 && lane_exec)
        atomicAdd((unsigned long long*)(F64 ? &ctl->steps64 : &ctl->steps32), (unsigned long long)lane_exec);
    if (lane == 0 && lane_rays) {
        atomicAdd((unsigned long long*)&ctl->steps, (unsigned long long)lane_steps);
        atomicAdd((unsigned long long*)&ctl->rays, (unsigned long long)lane_rays);
        if (F64) atomicAdd((unsigned long long*)&ctl->rays64, (unsigned long long)lane_rays);
    }
}

// Set-up of an FP32 pass (PASS 0: every ray; 1: mixed, deferring hard rays):
// one thread per work item computes the camera ray, the initial momenta and
// constants (init_ray) and, in the mixed pass, whether FP32 is accurate
// enough for it (hard_ray).  A ray to trace gets its init entry and record
// index; a ray that fails to start gets its record here and counts as
// finished; a hard ray is appended to the FP64 list (one atomic per wave).
// Moving this out of the persistent trace kernel keeps its lane refill to two
// loads: with the set-up inline, every refill ran the set-up code for the
// whole wave.
template <int METRIC, int PASS>
__global__ __launch_bounds__(256) void setup_kernel(SceneConst<float> sf, const float2* __restrict__ halton, Chunk ch,
                                                    QueueCtl* __restrict__ ctl) {
    constexpr bool KERR = (METRIC >= 1);
    const int lane = __lane_id();
    const std::uint64_t n_items = std::uint64_t(ch.S) * ch.n_tiles * 64u;
    const std::uint64_t stride = std::uint64_t(gridDim.x) * 256u;
    std::uint64_t fails = 0;
    // every lane of a wave runs the same number of iterations (the bound is
    // rounded up to whole waves), so the ballot below sees all 64 lanes
    const std::uint64_t start = std::uint64_t(blockIdx.x) * 256u + threadIdx.x;
    const std::uint64_t first_lane = start - std::uint64_t(lane);
    for (std::uint64_t base = first_lane; base < n_items; base += stride) {
        const std::uint64_t item = base + std::uint64_t(lane);
        std::uint32_t out = kItemSkip, rec_idx = 0, gx = 0, gy = 0, s = 0, kind = 0;
        bool hard = false;
        if (item < n_items && item_to_ray(item, ch, sf.width, sf.height, rec_idx, gx, gy, s)) {
            float ox = 0.5f, oy = 0.5f;
            if (sf.spp > 1u) {
                const float2 hv = halton[s];
                ox = hv.x;
                oy = hv.y;
            }
            RayState<float> y{};
            const bool started =
                init_ray<float, float, KERR>(y, sf, gx, gy, ox, oy, PASS == 1 ? &hard : nullptr, &kind);
            if (!started) {
                Rec rc;
                rc.a = rc.b = rc.c = 0.0f;
                rc.meta = kTermInitFail;
                ch.rec[rec_idx] = rc;
                fails += 1;
                hard = false;
            } else if (!hard) {
                ch.init[item] = make_float4(y.v0, y.v1, y.v2, y.v3);
                out = rec_idx;
            }
        }
        if (PASS == 1) {
            // long hard rays at the front of the FP64 list, short ones from its
            // end backwards (the FP64 pass takes the front first)
            const bool front = hard && kind != 2u, back = hard && kind == 2u;
            const std::uint64_t hf = __ballot(front), hb = __ballot(back);
            const std::uint64_t below = (1ull << lane) - 1ull;
            if (hf) {
                const int leader = __ffsll((unsigned long long)hf) - 1;
                std::uint32_t slot = 0;
                if (lane == leader) slot = atomicAdd(&ctl->band_count, std::uint32_t(__popcll(hf)));
                slot = __shfl(slot, leader);
                if (front) ch.band[slot + __popcll(hf & below)] = rec_idx;
            }
            if (hb) {
                const int leader = __ffsll((unsigned long long)hb) - 1;
                std::uint32_t slot = 0;
                if (lane == leader) slot = atomicAdd(&ctl->band_back, std::uint32_t(__popcll(hb)));
                slot = __shfl(slot, leader);
                if (back) ch.band[ch.S * ch.P - 1u - (slot + __popcll(hb & below))] = rec_idx;
            }
            if (hard) ch.rec[rec_idx] = deferred_marker();
        }
        if (item < n_items) ch.item_rec[item] = out;
    }
    for (int off = 32; off > 0; off >>= 1) fails += shfl_u64(fails, lane ^ off);
    if (lane == 0 && fails) atomicAdd((unsigned long long*)&ctl->rays, (unsigned long long)fails);
}

// Mixed precision: escape records whose direction sits on a starfield cell
// boundary next to a star are handed to the FP64 pass (kept out of the trace
// kernel so its step loop stays register-lean).
__global__ __launch_bounds__(256) void audit_kernel(SceneConst<float> sf, Chunk ch, QueueCtl* __restrict__ ctl) {
    const std::uint32_t n = ch.S * ch.P;
    for (std::uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const Rec r = ch.rec[i];
        const std::uint32_t term = r.meta & 7u;
        if (term == kTermEscape && starfield_ambiguous(r.a, r.b, sf.tol > 0.0f ? kStarCellAdaptive : kStarCell)) {
            const std::uint32_t slot = atomicAdd(&ctl->band_count, 1u);
            ch.band[slot] = i;
            atomicSub((unsigned long long*)&ctl->rays, 1ull);
            atomicSub((unsigned long long*)&ctl->steps, (unsigned long long)(r.meta >> 4));
        }
    }
}


template <int METRIC>
__global__ __launch_bounds__(256) void shade_kernel(SceneConst<float> sc, const Rec* __restrict__ rec,
                                                    PixAcc* __restrict__ acc, std::uint32_t P, std::uint32_t S,
                                                    std::uint32_t first, std::uint32_t last, TraceOutputs out) {
    constexpr bool KERR = (METRIC == 1);
    const std::uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= P) return;
    float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f, h0 = 0.0f, h1 = 0.0f, h2 = 0.0f;
    std::uint32_t steps = 0, cls = 0;
    if (!first) {
        const PixAcc a = acc[pix];
        l0 = a.ldr[0]; l1 = a.ldr[1]; l2 = a.ldr[2];
        h0 = a.hdr[0]; h1 = a.hdr[1]; h2 = a.hdr[2];
        steps = a.steps;
        cls = a.cls;
    }
    const float ex = sc.exposure;
    for (std::uint32_t s = 0; s < S; ++s) {
        const Rec r = rec[std::size_t(s) * P + pix];
        const std::uint32_t term = r.meta & 7u;
        float c[3] = {0.0f, 0.0f, 0.0f};
        if (term == kTermEscape) shade_escape<float>(sc, r.a, r.b, c);
        else if (term == kTermDisc) shade_disc<float, KERR>(sc, r.a, r.b, r.c, c);
        h0 += c[0];
        h1 += c[1];
        h2 += c[2];
        l0 += aces1(c[0] * ex);
        l1 += aces1(c[1] * ex);
        l2 += aces1(c[2] * ex);
        steps += r.meta >> 4;
        if (first && s == 0) cls = class_of(term);
    }
    if (!last) {
        PixAcc a;
        a.ldr[0] = l0; a.ldr[1] = l1; a.ldr[2] = l2;
        a.hdr[0] = h0; a.hdr[1] = h1; a.hdr[2] = h2;
        a.steps = steps;
        a.cls = cls;
        acc[pix] = a;
        return;
    }
    const std::uint32_t spp = sc.spp;
    if (spp > 1u) {
        const float inv = 1.0f / float(spp);
        l0 *= inv; l1 *= inv; l2 *= inv;
        h0 *= inv; h1 *= inv; h2 *= inv;
    }
    if (sc.grid) {  // tone map of the mean radiance (render/metal/shaders/blit.metal:64-68)
        l0 = aces1(h0 * ex);
        l1 = aces1(h1 * ex);
        l2 = aces1(h2 * ex);
    }
    if (out.rgba) out.rgba[pix] = make_uchar4(to_u8(l0), to_u8(l1), to_u8(l2), 255);
    if (out.hdr) {
        out.hdr[3u * pix + 0u] = h0;
        out.hdr[3u * pix + 1u] = h1;
        out.hdr[3u * pix + 2u] = h2;
    }
    if (out.hdr16) {
        const __half2 ab = __floats2half2_rn(h0, h1);
        const __half2 cd = __floats2half2_rn(h2, 1.0f);
        out.hdr16[pix] = make_uint2(*reinterpret_cast<const std::uint32_t*>(&ab),
                                    *reinterpret_cast<const std::uint32_t*>(&cd));
    }
    if (out.cls) out.cls[pix] = std::uint8_t(cls);
    if (out.steps) out.steps[pix] = steps;
}

}  // namespace

hipError_t launch_trace_chunk(const SceneConst<float>& sf, const SceneConst<double>& sd, const float2* halton,
                              const Chunk& ch, QueueCtl* ctl, std::uint32_t mode, const LaunchShape& shape,
                              hipStream_t stream, const hipEvent_t* ev, hipStream_t tail, hipEvent_t join) {
    auto mark = [&](int i) {
        if (ev) (void)hipEventRecord(ev[i], stream);
    };
    const dim3 block(256);
    const bool kerr = sf.metric == SING_METRIC_KERR;
    const bool adapt = kerr && sf.tol > 0.0f;
    // Step batches (RK4 steps, or error-controlled attempts, between lane
    // refills) from launch-shape sweeps on MI355X: the FP32 passes run best
    // at 64 (4K Kerr frame 1 % and the error-controlled 1080p frame 5 % faster
    // than at 32), the error-controlled full FP64 pass at 64 (2 %), the
    // fixed-step full FP64 pass at 32.  The mixed form's FP64 list mixes short
    // captured rays (~70 steps) with long near-polar ones (~600) and refills
    // more often: every quarter of the step batch (at least 8, a multiple of
    // 8), i.e. 16 by default (64 uneven scenes 3.6 % faster than at 8, the 4K
    // frame 0.1 % slower), and every 8 attempts with error control by default
    // (that frame's list 2 % faster than at 16).
    const std::uint32_t sb = shape.step_batch ? shape.step_batch : 64u;
    const std::uint32_t sb64 = shape.step_batch ? shape.step_batch : (adapt ? 64u : 32u);
    const std::uint32_t sb_list = (adapt && !shape.step_batch) ? 8u : std::max(8u, (sb / 4u) & ~7u);
#define SING_TRACE(P, G, SB)                                                                              \
    do {                                                                                                  \
        if (adapt)                                                                                        \
            hipLaunchKernelGGL((trace_kernel<2, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, SB); \
        else if (kerr && sf.refine)                                                                       \
            hipLaunchKernelGGL((trace_kernel<3, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, SB); \
        else if (kerr)                                                                                    \
            hipLaunchKernelGGL((trace_kernel<1, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, SB); \
        else                                                                                              \
            hipLaunchKernelGGL((trace_kernel<0, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, SB); \
    } while (0)
    const std::uint64_t n_items = std::uint64_t(ch.S) * ch.n_tiles * 64u;
    const dim3 setup_grid(std::uint32_t(std::min<std::uint64_t>((n_items + 255u) / 256u, shape.blocks * 8ull)));
#define SING_SETUP(P)                                                                                          \
    do {                                                                                                       \
        if (kerr)                                                                                              \
            hipLaunchKernelGGL((setup_kernel<1, P>), setup_grid, block, 0, stream, sf, halton, ch, ctl);        \
        else                                                                                                   \
            hipLaunchKernelGGL((setup_kernel<0, P>), setup_grid, block, 0, stream, sf, halton, ch, ctl);        \
    } while (0)
    if (mode == SING_PREC_FP32) {
        SING_SETUP(0);
        mark(0);
        SING_TRACE(0, shape.blocks, sb);
        mark(1);
        mark(2);
        mark(3);
    } else if (mode == SING_PREC_FP64) {
        mark(0);
        mark(1);
        mark(2);
        SING_TRACE(2, shape.blocks64, sb64);
        mark(3);
    } else {
        SING_SETUP(1);
        mark(0);
        SING_TRACE(1, shape.blocks, sb);
        mark(1);
        if (sf.flags & SING_FLAG_STARFIELD_ON)
            hipLaunchKernelGGL(audit_kernel, dim3(shape.blocks), block, 0, stream, sf, ch, ctl);
        if (tail && tail != stream) {
            hipError_t e = hipEventRecord(join, stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(tail, join, 0);
            if (e != hipSuccess) return e;
            stream = tail;
        }
        mark(2);
        SING_TRACE(3, shape.blocks64, sb_list);
        mark(3);
    }
#undef SING_TRACE
#undef SING_SETUP
    return hipGetLastError();
}

hipError_t launch_shade_chunk(const SceneConst<float>& sf, const Chunk& ch, PixAcc* acc, bool first, bool last,
                              const TraceOutputs& out, hipStream_t stream) {
    const dim3 grid((ch.P + 255u) / 256u), block(256);
    if (sf.metric == SING_METRIC_KERR)
        hipLaunchKernelGGL(shade_kernel<1>, grid, block, 0, stream, sf, ch.rec, acc, ch.P, ch.S, first ? 1u : 0u,
                           last ? 1u : 0u, out);
    else
        hipLaunchKernelGGL(shade_kernel<0>, grid, block, 0, stream, sf, ch.rec, acc, ch.P, ch.S, first ? 1u : 0u,
                           last ? 1u : 0u, out);
    return hipGetLastError();
}

}  // namespace sing
