// singularity_amd/csrc/batch.hip — batched render: many scenes of uneven
// sizes and sample counts in one persistent trace launch.
//
// Each wavefront works on one scene at a time, so the scene's constants are
// wave-uniform (scalar loads from the scene table, like kernel arguments in the
// single-scene kernel) and the per-ray code is exactly the single-scene one
// (trace_core.hpp).  Idle lanes claim rays from their scene's queue with one
// atomic per wave; when the scene runs dry the wave finishes its live rays and
// then moves to the next scene with work left, so small scenes never leave
// the GPU idle and big scenes are shared by many waves.  Termination records
// of the whole group stay in HBM (16 B per ray; the 64-scene workload is 23 GB),
// then one shade launch covers every pixel of every scene in the group.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/sing_abi.h"
#include "batch.hpp"
#include "context.hpp"
#include "trace_core.hpp"

namespace sing {

namespace {

using namespace core;

__device__ __forceinline__ std::uint64_t load_relaxed(const std::uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ std::uint32_t load_relaxed(const std::uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A wave-uniform 64-bit value in scalar registers.
__device__ __forceinline__ std::uint64_t uniform_u64(std::uint64_t v) {
    return (std::uint64_t(__builtin_amdgcn_readfirstlane(std::uint32_t(v >> 32))) << 32) |
           std::uint64_t(__builtin_amdgcn_readfirstlane(std::uint32_t(v)));
}
__device__ __forceinline__ std::uint64_t item_limit(std::uint64_t n_items, std::uint32_t sample_div) {
    if (sample_div <= 1u) return n_items;
    const std::uint64_t part = ((n_items / sample_div + 63u) / 64u) * 64u;
    return part < n_items ? part : n_items;
}

// Index of the last base <= v among bases[first .. first + count).
__device__ __forceinline__ std::uint32_t find_scene(const std::uint64_t* bases, std::uint32_t first,
                                                    std::uint32_t count, std::uint64_t v) {
    std::uint32_t lo = first, hi = first + count;  // answer in [lo, hi)
    while (hi - lo > 1u) {
        const std::uint32_t mid = (lo + hi) >> 1;
        if (bases[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void batch_reset_kernel(SceneQueue* __restrict__ q, std::uint32_t first,
                                                          std::uint32_t count, QueueCtl* __restrict__ ctl,
                                                          std::uint32_t reset_totals) {
    const std::uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < count) {
        SceneQueue z{};
        q[first + i] = z;
    }
    if (i == 0) {
        ctl->next_item = 0;
        ctl->next_band = 0;
        ctl->band_count = 0;
        if (reset_totals) {
            ctl->steps = 0;
            ctl->rays = 0;
            ctl->rays64 = 0;
            ctl->rejected = 0;
        }
    }
}

// METRIC 0: Schwarzschild; 1: Kerr; 2: Kerr with error-controlled steps.
// PASS as trace_kernel: 0 FP32, 1 FP32 with deferral, 2 FP64, 3 FP64 over the
// scenes' deferred lists.  Scenes [first, first + count) of the table.
// The scene table is its own read-only, non-aliased parameter so that, with a
// wave-uniform scene index, its fields are scalar loads (as kernel arguments are).
// The fixed-step Kerr FP32 passes are held to 80 VGPRs (6 waves per SIMD, as
// the single-scene trace): the inline ray set-up would otherwise take 104 (5
// waves); what that spills is reloaded in the claim code, once per step batch,
// not in the step loop.  The fixed-step Kerr FP64 passes are held to 168
// (3 waves per SIMD).
template <int METRIC, int PASS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(METRIC != 1 ? 1 : PASS < 2 ? 6 : 3)))
void trace_batch_kernel(const BatchScene* __restrict__ scenes, BatchLaunch L,
                                                          std::uint32_t first, std::uint32_t count,
                                                          std::uint32_t step_batch, std::uint32_t sample_div,
                                                          BatchClaim claim) {
    constexpr bool KERR = (METRIC >= 1);
    constexpr bool ADAPT = (METRIC == 2);
    constexpr bool F64 = (PASS >= 2);
    using S = typename std::conditional<F64, double, float>::type;
    const int lane = __lane_id();
    const std::uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
    std::uint32_t sid = __builtin_amdgcn_readfirstlane(first + wave % count);

    bool dry = false, need = true, live = false;
    std::uint32_t rec_idx = 0;  // relative to the scene's record base
    std::uint64_t lane_steps = 0, lane_rays = 0;
    std::uint32_t lane_rej = 0;
    RayState<S> y{};
    // the wave's private run of the current scene's work items (queue_chunk > 0)
    std::uint64_t pnext = 0, pend = 0;

    for (;;) {
        const BatchScene& bs = scenes[sid];
        SceneQueue& q = L.queues[sid];
        // (1) claim rays of the current scene for idle lanes: when at least
        // rays_per_refill lanes are idle (or none is live), in up to two rounds
        // (lanes whose item needed no trace, or that the private run could not
        // cover, go round again)
        if (!dry) {
            std::uint64_t want = __ballot(need);
            if (want && (std::uint32_t(__popcll(want)) >= claim.rays_per_refill || __ballot(live) == 0ull)) {
                const std::uint64_t n_front = (PASS == 3) ? std::uint64_t(q.band_count) : 0u;
                const std::uint64_t n_items =
                    (PASS == 3) ? n_front + q.band_back : item_limit(bs.n_items, sample_div);
                std::uint64_t* counter = (PASS == 3) ? &q.next_band : &q.next_item;
                for (int round = 0; round < 2 && want != 0ull && !dry; ++round) {
                    const std::uint32_t k = std::uint32_t(__popcll(want));
                    const int leader = __ffsll((unsigned long long)want) - 1;
                    std::uint64_t base;
                    std::uint32_t avail;
                    if (claim.queue_chunk <= 1u) {
                        base = 0;
                        if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)k);
                        base = uniform_u64(shfl_u64(base, leader));
                        avail = k;
                    } else {
                        if (pnext >= pend) {
                            std::uint64_t b = 0;
                            if (lane == leader)
                                b = atomicAdd((unsigned long long*)counter, (unsigned long long)claim.queue_chunk);
                            b = uniform_u64(shfl_u64(b, leader));
                            pnext = b;
                            pend = b + claim.queue_chunk;
                        }
                        const std::uint64_t left = pend - pnext;
                        avail = left < k ? std::uint32_t(left) : k;
                        base = pnext;
                        pnext += avail;
                    }
                    bool got_dry = false;
                    const std::uint32_t rank = std::uint32_t(__popcll(want & ((1ull << lane) - 1ull)));
                    if (need && rank < avail) {
                        const std::uint64_t item = base + rank;
                        if (item >= n_items) {
                            got_dry = true;
                        } else {
                            const std::uint32_t W = bs.sf.width, H = bs.sf.height;
                            std::uint32_t gx, gy, s;
                            bool ok;
                            if (PASS == 3) {
                                // the scene's FP64 list: long rays at the front, short
                                // ones at the end of its P * spp slots, front first
                                rec_idx = L.band[bs.rec_base + (item < n_front ? item
                                                                                 : std::uint64_t(bs.P) * bs.spp - 1u -
                                                                                       (item - n_front))];
                                const std::uint32_t sl = rec_idx / bs.P, pix = rec_idx - sl * bs.P;
                                gx = pix % W;
                                gy = pix / W;
                                s = sl;
                                ok = true;
                            } else {
                                Chunk ch{};
                                ch.s0 = 0;
                                ch.S = bs.spp;
                                ch.P = bs.P;
                                ch.tiles_x = bs.tiles_x;
                                ch.n_tiles = bs.n_tiles;
                                ok = item_to_ray(item, ch, W, H, rec_idx, gx, gy, s);
                            }
                            if (ok) {
                                float ox = 0.5f, oy = 0.5f;
                                if (bs.spp > 1u) {
                                    const float2 hv = L.halton[s];
                                    ox = hv.x;
                                    oy = hv.y;
                                }
                                bool started, deferred = false;
                                std::uint32_t kind = 0;
                                if (F64) {
                                    started = init_ray<double, S, KERR>(y, bs.sd, gx, gy, double(ox), double(oy),
                                                                        (bool*)nullptr);
                                } else if (PASS == 1) {
                                    bool hard = false;
                                    started = init_ray<float, S, KERR>(y, bs.sf, gx, gy, ox, oy, &hard, &kind);
                                    deferred = started && hard;
                                } else {
                                    started = init_ray<float, S, KERR>(y, bs.sf, gx, gy, ox, oy, (bool*)nullptr);
                                }
                                if (deferred) {
                                    if (kind == 2u) {  // short: from the end of the list
                                        const std::uint32_t slot = atomicAdd(&q.band_back, 1u);
                                        L.band[bs.rec_base + std::uint64_t(bs.P) * bs.spp - 1u - slot] = rec_idx;
                                    } else {
                                        const std::uint32_t slot = atomicAdd(&q.band_count, 1u);
                                        L.band[bs.rec_base + slot] = rec_idx;
                                    }
                                    L.rec[bs.rec_base + rec_idx] = deferred_marker();
                                } else if (!started) {
                                    Rec rc;
                                    rc.a = rc.b = rc.c = 0.0f;
                                    rc.meta = kTermInitFail;
                                    L.rec[bs.rec_base + rec_idx] = rc;
                                    lane_rays += 1;
                                } else {
                                    need = false;
                                    live = true;
                                }
                            }
                        }
                    }
                    if (__ballot(got_dry)) dry = true;
                    want = __ballot(need);
                }
            }
        }
        // (2) a dry scene with no live ray left: move to another scene with work
        if (dry && __ballot(live) == 0ull) {
            bool found = false;
            for (std::uint32_t j = 1; j <= count; ++j) {
                const std::uint32_t s2 = first + (sid - first + j) % count;
                const SceneQueue& q2 = L.queues[s2];
                const std::uint64_t lim =
                    (PASS == 3) ? std::uint64_t(load_relaxed(&q2.band_count)) + load_relaxed(&q2.band_back)
                                : item_limit(scenes[s2].n_items, sample_div);
                const std::uint64_t cur = load_relaxed((PASS == 3) ? &q2.next_band : &q2.next_item);
                if (__builtin_amdgcn_readfirstlane(cur < lim ? 1u : 0u)) {
                    sid = __builtin_amdgcn_readfirstlane(s2);
                    found = true;
                    break;
                }
            }
            if (!found) break;
            dry = false;
            need = true;
            pnext = pend = 0;
            continue;
        }

        // (3) integrate a batch of steps.  The step loop reads the scene through a
        // local copy: with the table reference, the compiler (at its SGPR limit)
        // re-loaded a constant from the table on every step and waited for it.
        std::uint32_t term;
        if constexpr (F64) {
            const SceneConst<double> scl = bs.sd;
            term = run_steps<double, S, KERR, false, ADAPT>(y, scl, live, step_batch, &lane_rej);
        } else {
            const SceneConst<float> scl = bs.sf;
            term = run_steps<float, S, KERR, PASS == 1, ADAPT>(y, scl, live, step_batch, &lane_rej);
        }

        // (4) emit finished rays; in the mixed pass, ambiguous outcomes go to FP64
        if (term != kTermLive) {
            bool defer = false;
            if (PASS == 1) defer = needs_fp64(y, term);
            if (defer) {
                const std::uint32_t slot = atomicAdd(&q.band_count, 1u);
                L.band[bs.rec_base + slot] = rec_idx;
                L.rec[bs.rec_base + rec_idx] = deferred_marker();
            } else {
                L.rec[bs.rec_base + rec_idx] = make_rec<S, KERR, ADAPT>(y, term);
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
        if (ADAPT) rej += shfl_u64(rej, lane ^ off);
    }
    if (lane == 0) {
        if (ADAPT && rej) atomicAdd((unsigned long long*)&L.ctl->rejected, (unsigned long long)rej);
        if (lane_rays) {
            atomicAdd((unsigned long long*)&L.ctl->steps, (unsigned long long)lane_steps);
            atomicAdd((unsigned long long*)&L.ctl->rays, (unsigned long long)lane_rays);
            if (F64) atomicAdd((unsigned long long*)&L.ctl->rays64, (unsigned long long)lane_rays);
        }
    }
}

// Mixed precision: escape records on a starfield cell edge next to a star go to
// their scene's FP64 list (as audit_kernel does for one scene).
__global__ __launch_bounds__(256) void audit_batch_kernel(BatchLaunch L, std::uint32_t first, std::uint32_t count,
                                                          std::uint64_t n_rec) {
    const std::uint64_t base0 = L.rec_bases[first];
    for (std::uint64_t i = std::uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n_rec;
         i += std::uint64_t(gridDim.x) * 256u) {
        const Rec r = L.rec[base0 + i];
        if ((r.meta & 7u) != kTermEscape) continue;
        const std::uint32_t sid = find_scene(L.rec_bases, first, count, base0 + i);
        const BatchScene& bs = L.scenes[sid];
        if (!(bs.sf.flags & SING_FLAG_STARFIELD_ON)) continue;
        if (starfield_ambiguous(r.a, r.b, bs.sf.tol > 0.0f ? kStarCellAdaptive : kStarCell)) {
            const std::uint32_t slot = atomicAdd(&L.queues[sid].band_count, 1u);
            L.band[bs.rec_base + slot] = std::uint32_t(base0 + i - bs.rec_base);
            atomicSub((unsigned long long*)&L.ctl->rays, 1ull);
            atomicSub((unsigned long long*)&L.ctl->steps, (unsigned long long)(r.meta >> 4));
        }
    }
}

// One thread per pixel of the group: the scene's samples in order, per-sample
// ACES of the exposed radiance, mean, sRGB (geodesic_kernel.cu:462-510).
__global__ __launch_bounds__(256) void shade_batch_kernel(BatchLaunch L, std::uint32_t first, std::uint32_t count,
                                                          std::uint64_t n_pix, BatchOutputs out) {
    const std::uint64_t g = std::uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (g >= n_pix) return;
    const std::uint64_t gp = L.pix_bases[first] + g;
    const std::uint32_t sid = find_scene(L.pix_bases, first, count, gp);
    const BatchScene& bs = L.scenes[sid];
    const SceneConst<float>& sc = bs.sf;
    const std::uint32_t pix = std::uint32_t(gp - bs.pix_base);
    const bool kerr = sc.metric == SING_METRIC_KERR;
    float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f, h0 = 0.0f, h1 = 0.0f, h2 = 0.0f;
    std::uint32_t cls = 0;
    const float ex = sc.exposure;
    const Rec* rec = L.rec + bs.rec_base;
    for (std::uint32_t s = 0; s < bs.spp; ++s) {
        const Rec r = rec[std::size_t(s) * bs.P + pix];
        const std::uint32_t term = r.meta & 7u;
        float c[3] = {0.0f, 0.0f, 0.0f};
        if (term == kTermEscape) shade_escape<float>(sc, r.a, r.b, c);
        else if (term == kTermDisc) {
            if (kerr) shade_disc<float, true>(sc, r.a, r.b, r.c, c);
            else shade_disc<float, false>(sc, r.a, r.b, r.c, c);
        }
        h0 += c[0];
        h1 += c[1];
        h2 += c[2];
        l0 += aces1(c[0] * ex);
        l1 += aces1(c[1] * ex);
        l2 += aces1(c[2] * ex);
        if (s == 0) cls = class_of(term);
    }
    if (bs.spp > 1u) {
        const float inv = 1.0f / float(bs.spp);
        l0 *= inv; l1 *= inv; l2 *= inv;
        h0 *= inv; h1 *= inv; h2 *= inv;
    }
    if (out.rgba) out.rgba[gp] = make_uchar4(to_u8(l0), to_u8(l1), to_u8(l2), 255);
    if (out.hdr) {
        out.hdr[3u * gp + 0u] = h0;
        out.hdr[3u * gp + 1u] = h1;
        out.hdr[3u * gp + 2u] = h2;
    }
    if (out.cls) out.cls[gp] = std::uint8_t(cls);
}

}  // namespace

hipError_t launch_batch_reset(const BatchLaunch& L, const BatchGroup& g, bool reset_totals, hipStream_t stream) {
    const std::uint32_t n = std::max(g.count, 1u);
    hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, L.queues, g.first, g.count,
                       L.ctl, reset_totals ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_batch_trace(const BatchLaunch& L, const BatchGroup& g, std::uint32_t mode, bool adaptive,
                              const LaunchShape& shape, const BatchClaim& claim, std::uint32_t sample_div,
                              bool starfield_any, hipStream_t stream) {
    const dim3 block(256);
    const std::uint32_t sb = shape.step_batch ? shape.step_batch : 32u;
    const std::uint32_t sb_list = std::max(8u, (sb / 4u) & ~7u);  // the FP64 list pass, as in trace.hip
    const std::uint32_t n_kerr = g.count - g.n_schw, k0 = g.first + g.n_schw;
#define SING_BTRACE(P, G, SB)                                                                                   \
    do {                                                                                                        \
        if (g.n_schw)                                                                                           \
            hipLaunchKernelGGL((trace_batch_kernel<0, P>), dim3(G), block, 0, stream, L.scenes, L, g.first,       \
                               g.n_schw, SB, sample_div, claim);                                                \
        if (n_kerr && adaptive)                                                                                 \
            hipLaunchKernelGGL((trace_batch_kernel<2, P>), dim3(G), block, 0, stream, L.scenes, L, k0, n_kerr, SB, \
                               sample_div, claim);                                                              \
        else if (n_kerr)                                                                                        \
            hipLaunchKernelGGL((trace_batch_kernel<1, P>), dim3(G), block, 0, stream, L.scenes, L, k0, n_kerr, SB, \
                               sample_div, claim);                                                              \
    } while (0)
    if (mode == SING_PREC_FP32) {
        SING_BTRACE(0, shape.blocks, sb);
    } else if (mode == SING_PREC_FP64) {
        SING_BTRACE(2, shape.blocks64, sb);
    } else {
        SING_BTRACE(1, shape.blocks, sb);
        if (starfield_any && sample_div <= 1u)
            hipLaunchKernelGGL(audit_batch_kernel, dim3(shape.blocks), block, 0, stream, L, g.first, g.count, g.rays);
        SING_BTRACE(3, shape.blocks64, sb_list);
    }
#undef SING_BTRACE
    return hipGetLastError();
}

hipError_t launch_batch_shade(const BatchLaunch& L, const BatchGroup& g, const BatchOutputs& out,
                              hipStream_t stream) {
    const std::uint64_t blocks = (g.pixels + 255u) / 256u;
    hipLaunchKernelGGL(shade_batch_kernel, dim3(std::uint32_t(blocks)), dim3(256), 0, stream, L, g.first, g.count,
                       g.pixels, out);
    return hipGetLastError();
}

}  // namespace sing

// ============================================================= host side
namespace {

// Termination records of one group (16 B per ray): groups are split so this fits.
constexpr std::size_t kBatchRecordBudget = std::size_t(32) << 30;

template <class T> int realloc_dev(SingContext* c, T** p, std::size_t& cap, std::size_t need, const char* what) {
    if (need <= cap && *p) return SING_OK;
    (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<std::size_t>(need, 1) * sizeof(T));
    if (e != hipSuccess) return c->fail(SING_ERR_NOMEM, "hipMalloc of batch %s (%zu x %zu B) failed: %s", what, need,
                                        sizeof(T), hipGetErrorString(e));
    cap = need;
    c->drop_batch_graph();  // a recorded graph refers to the old buffer
    return SING_OK;
}

sing::LaunchShape shape_of(const SingContext* c, std::uint32_t step_batch, std::uint32_t wpc) {
    SingRenderOpts ro{};
    ro.step_batch = step_batch;
    ro.waves_per_cu = wpc;
    return c->launch_shape(ro);
}

std::uint64_t tune_key(bool has_schw, bool has_kerr, std::uint32_t precision, bool adaptive) {
    return (has_schw ? 1u : 0u) | (has_kerr ? 2u : 0u) | (std::uint64_t(precision) << 2) | (adaptive ? 64u : 0u);
}

int run_groups(SingContext* c, const sing::BatchLaunch& L, const std::vector<sing::BatchGroup>& groups,
               std::uint32_t precision, bool adaptive, const sing::LaunchShape& shape, const sing::BatchClaim& claim,
               bool starfield_any, const sing::BatchOutputs& out) {
    hipError_t e = hipSuccess;
    for (std::size_t k = 0; k < groups.size() && e == hipSuccess; ++k) {
        const sing::BatchGroup& g = groups[k];
        e = sing::launch_batch_reset(L, g, k == 0, c->stream);
        if (e == hipSuccess)
            e = sing::launch_batch_trace(L, g, precision, adaptive, shape, claim, 1u, starfield_any, c->stream);
        if (e == hipSuccess) e = sing::launch_batch_shade(L, g, out, c->stream);
    }
    return e == hipSuccess ? SING_OK : c->hip_fail(e, "batch launch");
}

}  // namespace

void SingContext::drop_batch_graph() {
    if (b_graph_exec) (void)hipGraphExecDestroy(b_graph_exec);
    if (b_graph) (void)hipGraphDestroy(b_graph);
    b_graph_exec = nullptr;
    b_graph = nullptr;
    b_graph_key.clear();
}

void SingContext::free_batch() {
    drop_batch_graph();
    (void)hipFree(d_bscenes);
    (void)hipFree(d_brec_bases);
    (void)hipFree(d_bpix_bases);
    (void)hipFree(d_bqueues);
    (void)hipFree(d_brec);
    (void)hipFree(d_bband);
    (void)hipFree(d_brgba);
    (void)hipFree(d_bhdr);
    (void)hipFree(d_bcls);
    d_bscenes = nullptr;
    d_brec_bases = d_bpix_bases = nullptr;
    d_bqueues = nullptr;
    d_brec = nullptr;
    d_bband = nullptr;
    d_brgba = nullptr;
    d_bhdr = nullptr;
    d_bcls = nullptr;
    b_scene_cap = b_rb_cap = b_pb_cap = b_q_cap = b_rec_cap = b_band_cap = b_rgba_cap = b_hdr_cap = b_cls_cap = 0;
    b_tuned.clear();
}

extern "C" int sing_render_batch(SingContext* c, const SingBatchItem* items, uint32_t n, const SingBatchOpts* opts_in,
                                 SingBatchStats* stats) {
    if (!c) return SING_ERR_ARG;
    c->complete_pending();
    if (!items || n == 0) return c->fail(SING_ERR_ARG, "sing_render_batch: no scenes");
    if (n > 65536u) return c->fail(SING_ERR_ARG, "sing_render_batch: at most 65536 scenes per batch");
    SingBatchOpts o{};
    if (opts_in) o = *opts_in;
    if (o.precision > SING_PREC_MIXED)
        return c->fail(SING_ERR_ARG, "sing_render_batch: unknown precision %u", o.precision);
    if (!(o.adapt_tol >= 0.0f && o.adapt_tol <= 0.1f))
        return c->fail(SING_ERR_ARG, "sing_render_batch: adapt_tol must be in [0, 0.1]");
    if (o.use_graph > 1u || o.autotune > 1u) return c->fail(SING_ERR_ARG, "sing_render_batch: flags must be 0 or 1");
    if (o.rays_per_refill > 64u) return c->fail(SING_ERR_ARG, "sing_render_batch: rays_per_refill must be in [0, 64]");
    if (o.queue_chunk > (1u << 20))
        return c->fail(SING_ERR_ARG, "sing_render_batch: queue_chunk must be at most 2^20");

    // scene order: Schwarzschild first, then Kerr (stable), so each metric is one
    // contiguous run per group and one kernel instantiation traces it
    std::vector<std::uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    for (std::uint32_t i = 0; i < n; ++i) {
        const SingUniforms* u = items[i].uniforms;
        if (!u) return c->fail(SING_ERR_ARG, "sing_render_batch: scene %u has no uniforms", i);
        if (u->width < 1 || u->height < 1 || u->width > 32768 || u->height > 32768)
            return c->fail(SING_ERR_ARG, "sing_render_batch: scene %u has bad size %ux%u", i, u->width, u->height);
        if (u->metric_type > SING_METRIC_KERR)
            return c->fail(SING_ERR_ARG, "sing_render_batch: scene %u has unknown metric %u", i, u->metric_type);
    }
    std::stable_partition(order.begin(), order.end(),
                          [&](std::uint32_t i) { return items[i].uniforms->metric_type == SING_METRIC_SCHWARZSCHILD; });

    SingRenderOpts ro{};
    ro.precision = o.precision;
    ro.adapt_tol = o.adapt_tol;
    ro.turb_seed = o.turb_seed;
    std::vector<sing::BatchScene> tab(n);
    std::vector<std::uint64_t> rec_bases(n), pix_bases(n);
    std::vector<sing::BatchGroup> groups;
    const std::size_t budget =
        o.record_budget_mib ? std::size_t(o.record_budget_mib) << 20 : kBatchRecordBudget;
    const std::uint64_t ray_budget = budget / sizeof(sing::Rec);
    std::uint64_t pix_total = 0, ray_total = 0, max_group_rays = 0;
    bool want_rgba = false, want_hdr = false, want_cls = false, starfield_any = false, has_schw = false,
         has_kerr = false;
    sing::BatchGroup cur{};
    for (std::uint32_t k = 0; k < n; ++k) {
        const SingBatchItem& it = items[order[k]];
        const SingUniforms& u = *it.uniforms;
        sing::BatchScene& b = tab[k];
        b.sf = sing::make_scene_const<float>(u, ro);
        b.sd = sing::make_scene_const<double>(u, ro);
        if (u.metric_type != SING_METRIC_KERR) b.sf.tol = b.sd.tol = 0.0;  // error control is Kerr-only
        b.P = u.width * u.height;
        b.spp = b.sf.spp;
        b.tiles_x = (u.width + 7u) / 8u;
        b.n_tiles = b.tiles_x * ((u.height + 7u) / 8u);
        b.n_items = std::uint64_t(b.spp) * b.n_tiles * 64u;
        const std::uint64_t rays = std::uint64_t(b.P) * b.spp;
        if (rays >= (std::uint64_t(1) << 31) || rays > ray_budget)
            return c->fail(SING_ERR_ARG,
                           "sing_render_batch: scene %u has %llu rays (limit 2^31 per scene and the record budget)",
                           order[k], (unsigned long long)rays);
        if (cur.count > 0 && cur.rays + rays > ray_budget) {
            groups.push_back(cur);
            cur = sing::BatchGroup{};
            cur.first = k;
            cur.pix_base = pix_total;
        }
        if (cur.count == 0) {
            cur.first = k;
            cur.pix_base = pix_total;
        }
        b.rec_base = cur.rays;
        b.pix_base = pix_total;
        rec_bases[k] = b.rec_base;
        pix_bases[k] = b.pix_base;
        cur.count += 1;
        if (u.metric_type == SING_METRIC_SCHWARZSCHILD) cur.n_schw += 1;
        cur.rays += rays;
        cur.pixels += b.P;
        pix_total += b.P;
        ray_total += rays;
        want_rgba |= it.rgba != nullptr;
        want_hdr |= it.hdr != nullptr;
        want_cls |= it.cls != nullptr;
        starfield_any |= (u.flags & SING_FLAG_STARFIELD_ON) != 0;
        has_schw |= u.metric_type == SING_METRIC_SCHWARZSCHILD;
        has_kerr |= u.metric_type == SING_METRIC_KERR;
    }
    groups.push_back(cur);
    for (const auto& g : groups) max_group_rays = std::max(max_group_rays, g.rays);
    const bool adaptive = o.adapt_tol > 0.0f && has_kerr;

    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hip_fail(e, "hipSetDevice");
    int rc;
    if ((rc = realloc_dev(c, &c->d_bscenes, c->b_scene_cap, n, "scene table")) != SING_OK ||
        (rc = realloc_dev(c, &c->d_brec_bases, c->b_rb_cap, n, "record bases")) != SING_OK ||
        (rc = realloc_dev(c, &c->d_bpix_bases, c->b_pb_cap, n, "pixel bases")) != SING_OK ||
        (rc = realloc_dev(c, &c->d_bqueues, c->b_q_cap, n, "queues")) != SING_OK)
        return rc;
    if ((rc = realloc_dev(c, &c->d_brec, c->b_rec_cap, max_group_rays, "records")) != SING_OK) return rc;
    if (o.precision == SING_PREC_MIXED &&
        (rc = realloc_dev(c, &c->d_bband, c->b_band_cap, max_group_rays, "deferred lists")) != SING_OK)
        return rc;
    if (want_rgba && (rc = realloc_dev(c, &c->d_brgba, c->b_rgba_cap, pix_total, "RGBA8 outputs")) != SING_OK)
        return rc;
    if (want_hdr && (rc = realloc_dev(c, &c->d_bhdr, c->b_hdr_cap, 3 * pix_total, "HDR outputs")) != SING_OK)
        return rc;
    if (want_cls && (rc = realloc_dev(c, &c->d_bcls, c->b_cls_cap, pix_total, "class outputs")) != SING_OK)
        return rc;
    if ((rc = c->ensure_events(4)) != SING_OK) return rc;
    if ((e = hipMemcpy(c->d_bscenes, tab.data(), n * sizeof(sing::BatchScene), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->d_brec_bases, rec_bases.data(), n * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->d_bpix_bases, pix_bases.data(), n * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpy(batch table)");

    sing::BatchLaunch L{};
    L.scenes = c->d_bscenes;
    L.rec_bases = c->d_brec_bases;
    L.pix_bases = c->d_bpix_bases;
    L.queues = c->d_bqueues;
    L.halton = c->d_halton;
    L.rec = c->d_brec;
    L.band = c->d_bband;
    L.ctl = c->d_ctl;
    sing::BatchOutputs out{want_rgba ? c->d_brgba : nullptr, want_hdr ? c->d_bhdr : nullptr,
                           want_cls ? c->d_bcls : nullptr};

    // launch shape and claim policy: explicit, else tuned for this kind of batch, else the defaults:
    // 64 steps between refills, 24 waves per CU, a refill once 32 lanes are idle and private runs of
    // 256 items (what the tuner picks on 64 uneven 4K-class scenes; 1.54x faster there than refilling
    // every idle lane item by item), with shorter runs when there are few items per wave (a single
    // 128^2 scene ran twice as long with runs of 256)
    const std::uint64_t per_wave = ray_total / (std::uint64_t(std::max(c->num_cu, 1)) * 24u);
    const std::uint32_t default_qc = per_wave >= 1024u ? 256u : per_wave >= 256u ? 64u : 1u;
    const std::uint64_t tkey = tune_key(has_schw, has_kerr, o.precision, adaptive);
    sing::BatchTune tu{o.step_batch, o.waves_per_cu, o.rays_per_refill, o.queue_chunk};
    const bool explicit_rpr = o.rays_per_refill != 0u, explicit_qc = o.queue_chunk != 0u;
    for (const auto& kv : c->b_tuned)
        if (kv.first == tkey) {
            if (!tu.step_batch) tu.step_batch = kv.second.step_batch;
            if (!tu.waves_per_cu) tu.waves_per_cu = kv.second.waves_per_cu;
            if (!explicit_rpr) tu.rays_per_refill = kv.second.rays_per_refill;
            if (!explicit_qc) tu.queue_chunk = kv.second.queue_chunk;
        }
    if (o.autotune && (!o.step_batch || !o.waves_per_cu || !explicit_rpr || !explicit_qc)) {
        // coordinate search on the first 1/8 of the first group's work items:
        // waves per CU, step batch, rays per refill, queue chunk, each with the
        // best of the knobs before it.  A candidate replaces the best only when
        // it is at least 1 % faster: differences below that are the sample
        // launches' own spread, and following them picked shapes up to 3 %
        // slower on the full batch.
        auto time_shape = [&](const sing::BatchTune& t, float* ms) -> int {
            const sing::LaunchShape sh = shape_of(c, t.step_batch, t.waves_per_cu);
            const sing::BatchClaim cl{t.rays_per_refill, t.queue_chunk};
            hipError_t r = sing::launch_batch_reset(L, groups[0], true, c->stream);
            if (r == hipSuccess) r = hipEventRecord(c->events[0], c->stream);
            if (r == hipSuccess)
                r = sing::launch_batch_trace(L, groups[0], o.precision, adaptive, sh, cl, 8u, false, c->stream);
            if (r == hipSuccess) r = hipEventRecord(c->events[1], c->stream);
            if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
            if (r == hipSuccess) r = hipEventElapsedTime(ms, c->events[0], c->events[1]);
            return r == hipSuccess ? SING_OK : c->hip_fail(r, "batch autotune");
        };
        sing::BatchTune best_t{tu.step_batch ? tu.step_batch : 64u, tu.waves_per_cu ? tu.waves_per_cu : 24u,
                               tu.rays_per_refill ? tu.rays_per_refill : 32u,
                               tu.queue_chunk ? tu.queue_chunk : default_qc};
        float best = 0.0f, ms = 0.0f;
        if ((rc = time_shape(best_t, &best)) != SING_OK) return rc;  // also warms the code up
        if ((rc = time_shape(best_t, &best)) != SING_OK) return rc;
        auto sweep = [&](std::uint32_t sing::BatchTune::*field, std::initializer_list<std::uint32_t> vals) -> int {
            for (std::uint32_t v : vals) {
                sing::BatchTune t = best_t;
                t.*field = v;
                if (t.*field == best_t.*field) continue;
                const int r = time_shape(t, &ms);
                if (r != SING_OK) return r;
                if (ms < 0.99f * best) best = ms, best_t = t;
            }
            return SING_OK;
        };
        if (!o.waves_per_cu && (rc = sweep(&sing::BatchTune::waves_per_cu, {12u, 16u, 24u, 32u})) != SING_OK) return rc;
        if (!o.step_batch && (rc = sweep(&sing::BatchTune::step_batch, {8u, 16u, 32u, 64u})) != SING_OK) return rc;
        if (!explicit_rpr && (rc = sweep(&sing::BatchTune::rays_per_refill, {1u, 4u, 8u, 16u, 32u})) != SING_OK)
            return rc;
        if (!explicit_qc && (rc = sweep(&sing::BatchTune::queue_chunk, {1u, 64u, 256u})) != SING_OK) return rc;
        // the step batch again: fewer refills shift its best value
        if (!o.step_batch && (rc = sweep(&sing::BatchTune::step_batch, {8u, 16u, 32u, 64u})) != SING_OK) return rc;
        tu = best_t;
        bool stored = false;
        for (auto& kv : c->b_tuned)
            if (kv.first == tkey) kv.second = tu, stored = true;
        if (!stored) c->b_tuned.push_back({tkey, tu});
    }
    if (!tu.step_batch) tu.step_batch = 64u;
    if (!tu.waves_per_cu) tu.waves_per_cu = 24u;
    if (!tu.rays_per_refill) tu.rays_per_refill = 32u;
    if (!tu.queue_chunk) tu.queue_chunk = default_qc;
    const sing::LaunchShape shape = shape_of(c, tu.step_batch, tu.waves_per_cu);
    const std::uint32_t sb = shape.step_batch, wpc = tu.waves_per_cu;
    const sing::BatchClaim claim{tu.rays_per_refill, tu.queue_chunk};

    // graph key: everything the recorded launches depend on
    std::vector<std::uint64_t> key = {n, o.precision, adaptive ? 1u : 0u, sb, wpc, claim.rays_per_refill,
                                      claim.queue_chunk, starfield_any ? 1u : 0u,
                                      reinterpret_cast<std::uint64_t>(c->d_bscenes),
                                      reinterpret_cast<std::uint64_t>(c->d_brec),
                                      reinterpret_cast<std::uint64_t>(c->d_bband),
                                      reinterpret_cast<std::uint64_t>(out.rgba),
                                      reinterpret_cast<std::uint64_t>(out.hdr),
                                      reinterpret_cast<std::uint64_t>(out.cls)};
    for (std::uint32_t k = 0; k < n; ++k)
        key.push_back((std::uint64_t(tab[k].P) << 32) ^ (std::uint64_t(tab[k].spp) << 20) ^
                      (std::uint64_t(tab[k].sf.metric) << 62) ^ items[order[k]].uniforms->width);
    // the record-memory grouping (set by record_budget_mib) decides the recorded launches
    key.push_back(groups.size());
    for (const auto& g : groups) key.push_back((std::uint64_t(g.first) << 32) | g.count);
    std::uint32_t replayed = 0;
    (void)hipEventRecord(c->events[0], c->stream);
    if (o.use_graph) {
        if (c->b_graph_exec && c->b_graph_key == key) {
            replayed = 1;
        } else {
            c->drop_batch_graph();
            if ((e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return c->hip_fail(e, "hipStreamBeginCapture");
            const int r1 = run_groups(c, L, groups, o.precision, adaptive, shape, claim, starfield_any, out);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(c->stream, &g);
            if (r1 != SING_OK) {
                if (g) (void)hipGraphDestroy(g);
                return r1;
            }
            if (e != hipSuccess) return c->hip_fail(e, "hipStreamEndCapture");
            c->b_graph = g;
            if ((e = hipGraphInstantiate(&c->b_graph_exec, g, nullptr, nullptr, 0)) != hipSuccess) {
                c->drop_batch_graph();
                return c->hip_fail(e, "hipGraphInstantiate");
            }
            c->b_graph_key = key;
        }
        (void)hipEventRecord(c->events[0], c->stream);
        if ((e = hipGraphLaunch(c->b_graph_exec, c->stream)) != hipSuccess) return c->hip_fail(e, "hipGraphLaunch");
    } else {
        if ((rc = run_groups(c, L, groups, o.precision, adaptive, shape, claim, starfield_any, out)) != SING_OK)
            return rc;
    }
    (void)hipEventRecord(c->events[1], c->stream);
    sing::QueueCtl ctl{};
    if ((e = hipMemcpyAsync(&ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(batch stats)");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize(batch)");
    float kern = 0.0f;
    (void)hipEventElapsedTime(&kern, c->events[0], c->events[1]);

    // outputs to the callers' buffers
    (void)hipEventRecord(c->events[2], c->stream);
    for (std::uint32_t k = 0; k < n && e == hipSuccess; ++k) {
        const SingBatchItem& it = items[order[k]];
        const std::size_t P = tab[k].P, pb = tab[k].pix_base;
        if (it.rgba) e = hipMemcpyAsync(it.rgba, c->d_brgba + pb, P * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && it.hdr)
            e = hipMemcpyAsync(it.hdr, c->d_bhdr + 3 * pb, P * 12, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && it.cls) e = hipMemcpyAsync(it.cls, c->d_bcls + pb, P, hipMemcpyDeviceToHost, c->stream);
    }
    if (e != hipSuccess) return c->hip_fail(e, "hipMemcpyAsync(batch outputs)");
    (void)hipEventRecord(c->events[3], c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize(batch)");
    float copy_ms = 0.0f;
    (void)hipEventElapsedTime(&copy_ms, c->events[2], c->events[3]);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->kernel_ms = kern;
        stats->copy_ms = copy_ms;
        stats->rays = ctl.rays;
        stats->steps = ctl.steps;
        stats->fp64_rays = ctl.rays64;
        stats->rejected = ctl.rejected;
        stats->step_batch = sb;
        stats->waves_per_cu = wpc;
        stats->rays_per_refill = claim.rays_per_refill;
        stats->queue_chunk = claim.queue_chunk;
        stats->graph_replayed = replayed;
        stats->groups = std::uint32_t(groups.size());
    }
    if (ctl.rays != ray_total)
        return c->fail(SING_ERR_STATE, "sing_render_batch: traced %llu rays, expected %llu",
                       (unsigned long long)ctl.rays, (unsigned long long)ray_total);
    return SING_OK;
}
