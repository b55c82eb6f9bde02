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

#include <cstdint>
#include <type_traits>

#include "physics.hpp"
#include "trace.hpp"

namespace sing {

namespace {

template <class S> struct RayState {
    S r, th, ph;
    S v0, v1, v2, v3;  // Kerr: p_r, p_theta, E, L_z.  Schwarzschild: u^t, u^r, u^theta, u^phi
    S prev_r, prev_th, prev_uph;  // prev_uph: Schwarzschild u^phi; error-controlled Kerr: phi
    S prev_v0, prev_v1, hprev;    // error-controlled Kerr: p_r, p_theta before the step, its size
    S aux0, aux1;      // disc crossing radius; u^phi (Schwarzschild) or phi (error-controlled Kerr) at it
    S near_eq;         // mixed precision: min |theta - pi/2| over the steps (0 = ambiguous crossing)
    S hcur;            // error-controlled stepping: current step size
    std::uint32_t steps;
    std::uint32_t tries;  // error-controlled stepping: rejections of the current step
};

// Mixed-precision deferral thresholds.  Comparing FP32 and FP64 traces of the
// same rays, outcomes differ only for rays near a discontinuity of the
// discrete algorithm: a state within kGrazeTheta of the equator (the crossing
// test flips), a crossing within kEdgeRel of an annulus edge, or an escape
// within kStarCell cells of a starfield boundary next to a star.  Such rays
// are re-traced in FP64.
constexpr float kGrazeTheta = 2e-6f;
constexpr float kEdgeRel = 1e-5f;
constexpr float kStarCell = 1e-3f;
// With error-controlled steps FP32 and FP64 take different step sequences, so
// their exit directions differ by the integration error (~1e-5 rad), not by
// rounding: the starfield guard widens to 1e-2 cell.
constexpr float kStarCellAdaptive = 1e-2f;

__device__ __forceinline__ std::uint64_t shfl_u64(std::uint64_t v, int src) {
    const std::uint32_t lo = __shfl(std::uint32_t(v), src);
    const std::uint32_t hi = __shfl(std::uint32_t(v >> 32), src);
    return (std::uint64_t(hi) << 32) | lo;
}

// Termination checks in the reference's order (geodesic_kernel.cu:351-401)
// followed by one RK4 step.  Arithmetic in R, storage in S.  With ADAPT (Kerr)
// the step is one attempt of the error-controlled step; a rejected attempt
// leaves the state unchanged (the checks it repeats are idempotent) and
// counts in *rejected.
template <class R, class S, bool KERR, bool TRACK = false, bool ADAPT = false>
__device__ __forceinline__ std::uint32_t advance(RayState<S>& y, const SceneConst<R>& sc,
                                                 std::uint32_t* rejected = nullptr) {
    if (y.steps >= sc.max_steps) return kTermBudget;
    R r = R(y.r), th = R(y.th);
    if (!__builtin_isfinite(r) || r < sc.cut) return kTermHorizon;
    if (r > sc.escape_r) {
        if (ADAPT) {
            R the = th, phe = R(y.ph);
            if (y.steps > 0u) {
                const KerrK<R> k = kerr_k(sc, R(y.v2), R(y.v3));
                const KerrDense<R> d = kerr_dense(R(y.prev_r), R(y.prev_th), R(y.prev_uph), R(y.prev_v0),
                                                  R(y.prev_v1), r, th, R(y.ph), R(y.v0), R(y.v1), R(y.hprev), sc, k);
                kerr_dense_escape(d, sc.escape_r, the, phe);
            }
            y.aux0 = S(the);
            y.aux1 = S(phe);
        }
        return kTermEscape;
    }
    const R mid = R(1.5707963267948966192313216916397514);
    if (sc.flags & SING_FLAG_DISC_ON) {
        const R pth = R(y.prev_th);
        const R dp = pth - mid, dc = th - mid;
        if (dp * dc < R(0)) {
            const R frac = (mid - pth) / (th - pth);
            const R pr0 = R(y.prev_r);
            R rc = pr0 + frac * (r - pr0);
            R phc = R(0);
            if (ADAPT) {
                const KerrK<R> k = kerr_k(sc, R(y.v2), R(y.v3));
                const KerrDense<R> d = kerr_dense(pr0, pth, R(y.prev_uph), R(y.prev_v0), R(y.prev_v1), r, th, R(y.ph),
                                                  R(y.v0), R(y.v1), R(y.hprev), sc, k);
                kerr_dense_crossing(d, frac, rc, phc);
            }
            if (TRACK) {
                const R tol = R(kEdgeRel) * rc;
                const R d_in = rc - sc.r_in, d_out = rc - sc.r_out;
                if ((d_in < tol && d_in > -tol) || (d_out < tol && d_out > -tol)) y.near_eq = S(0);
            }
            if (rc >= sc.r_in && rc <= sc.r_out) {
                y.aux0 = S(rc);
                if (ADAPT) y.aux1 = S(phc);
                if (!KERR) {
                    const R pu = R(y.prev_uph);
                    y.aux1 = S(pu + frac * (R(y.v3) - pu));
                }
                return kTermDisc;
            }
        }
    }
    y.prev_th = y.th;
    y.prev_r = y.r;
    R ph = R(y.ph);
    if (KERR) {
        R pr = R(y.v0), pt = R(y.v1);
        const KerrK<R> k = kerr_k(sc, R(y.v2), R(y.v3));
        if (ADAPT) {
            R hc = R(y.hcur), hused;
            y.prev_uph = y.ph;
            y.prev_v0 = y.v0;
            y.prev_v1 = y.v1;
            const bool acc = kerr_adaptive_step(r, th, ph, pr, pt, hc, y.tries, sc.tol, sc, k, hused);
            y.hcur = S(hc);
            if (!acc) {
                *rejected += 1u;
                return kTermLive;
            }
            y.hprev = S(hused);
        } else {
            kerr_rk4(r, th, ph, pr, pt, sc.h, sc, k);
        }
        y.v0 = S(pr);
        y.v1 = S(pt);
    } else {
        y.prev_uph = y.v3;
        R ut = R(y.v0), ur = R(y.v1), uth = R(y.v2), uph = R(y.v3);
        schw_rk4(r, th, ph, ut, ur, uth, uph, sc.h, sc.rs);
        y.v0 = S(ut);
        y.v1 = S(ur);
        y.v2 = S(uth);
        y.v3 = S(uph);
    }
    y.r = S(r);
    y.th = S(th);
    y.ph = S(ph);
    y.steps += 1;
    if (TRACK) {
        const R d = th - mid;
        const R ad = d < R(0) ? -d : d;
        if (ad < R(y.near_eq)) y.near_eq = S(ad);
    }
    return kTermLive;
}

// Ray initialisation from the camera (geodesic_kernel.cu:229-242, 303-347).
// Returns false on a Kerr initialisation failure (no real root at the camera).
// With margin_out, also evaluates the photon-orbit margin of the ray.
template <class R, class S, bool KERR>
__device__ __forceinline__ bool init_ray(RayState<S>& y, const SceneConst<R>& sc, std::uint32_t gx, std::uint32_t gy,
                                         R ox, R oy, R* margin_out) {
    R ur, uth, uph;
    camera_ray(sc, gx, gy, ox, oy, ur, uth, uph);
    y.r = S(sc.cam_r);
    y.th = S(sc.cam_theta);
    y.ph = S(sc.cam_phi);
    y.prev_r = y.r;
    y.prev_th = y.th;
    y.steps = 0;
    y.near_eq = S(1e30);
    y.hcur = S(sc.h);
    y.tries = 0u;
    if (KERR) {
        const R Bq = R(2) * sc.g_tp * uph;
        const R Cq = sc.g_rr * ur * ur + sc.g_thth * uth * uth + sc.g_pp * uph * uph;
        const R disc = Bq * Bq - R(4) * sc.g_tt * Cq;
        if (disc < R(0)) return false;
        const R ut = (-Bq - dsqrt(disc)) / (R(2) * sc.g_tt);
        const R E = -(sc.g_tt * ut + sc.g_tp * uph);
        const R L = sc.g_tp * ut + sc.g_pp * uph;
        const R pt = sc.g_thth * uth;
        y.v0 = S(sc.g_rr * ur);
        y.v1 = S(pt);
        y.v2 = S(E);
        y.v3 = S(L);
        if (margin_out) {
            const R Q = pt * pt + sc.q_cos2 * (L * L * sc.q_inv_s2 - sc.a2 * E * E);
            const R rp = sc.M + dsqrt(dmax(sc.M * sc.M - sc.a2, R(0)));
            const R m = photon_margin<R>(L / E, Q / (E * E), sc.a, sc.M, rp);
            // no turning point: ingoing rays are captured (defer), outgoing escape directly
            *margin_out = (m < R(0)) ? (ur < R(0) ? R(-1) : R(1e30)) : m;
        }
    } else {
        const R cr = sc.cam_r;
        const R kin = (ur * ur) / sc.f_cam + uth * uth * cr * cr + uph * uph * cr * cr * sc.sin_th * sc.sin_th;
        const R ut = dsqrt(dmax(kin / sc.f_cam, R(0)));
        y.v0 = S(ut);
        y.v1 = S(ur);
        y.v2 = S(uth);
        y.v3 = S(uph);
        y.prev_uph = y.v3;
        if (margin_out) {
            const R E = sc.f_cam * ut;
            const R L2 = cr * cr * cr * cr * (uth * uth + sc.sin_th * sc.sin_th * uph * uph);
            const R m = photon_margin<R>(R(0), L2 / (E * E), R(0), sc.M, sc.twoM);
            *margin_out = (m < R(0)) ? (ur < R(0) ? R(-1) : R(1e30)) : m;
        }
    }
    return true;
}

template <class S, bool KERR, bool ADAPT = false>
__device__ __forceinline__ Rec make_rec(const RayState<S>& y, std::uint32_t term) {
    Rec rc;
    rc.meta = term | (y.steps << 4);
    if (term == kTermDisc) {
        rc.a = float(y.aux0);
        rc.b = ADAPT ? float(y.aux1) : float(y.ph);
        rc.c = KERR ? float(y.v3) : float(y.aux1);
    } else if (ADAPT && term == kTermEscape) {
        rc.a = float(y.aux0);  // dense-output direction at the escape radius
        rc.b = float(y.aux1);
        rc.c = 0.0f;
    } else {
        rc.a = float(y.th);
        rc.b = float(y.ph);
        rc.c = 0.0f;
    }
    return rc;
}

// Work item -> (record index, pixel, sample).  Items enumerate the chunk's
// samples sample-major, pixels in 8x8 tiles so a wave starts on a compact tile.
__device__ __forceinline__ bool item_to_ray(std::uint64_t item, const Chunk& ch, std::uint32_t W, std::uint32_t H,
                                            std::uint32_t& rec_idx, std::uint32_t& gx, std::uint32_t& gy,
                                            std::uint32_t& s) {
    const std::uint64_t per_s = std::uint64_t(ch.n_tiles) * 64u;
    const std::uint32_t sl = std::uint32_t(item / per_s);
    const std::uint32_t rem = std::uint32_t(item - std::uint64_t(sl) * per_s);
    const std::uint32_t tile = rem >> 6, w = rem & 63u;
    gx = (tile % ch.tiles_x) * 8u + (w & 7u);
    gy = (tile / ch.tiles_x) * 8u + (w >> 3);
    s = ch.s0 + sl;
    rec_idx = sl * ch.P + gy * W + gx;
    return gx < W && gy < H;
}

// METRIC 0: Schwarzschild; 1: Kerr; 2: Kerr with error-controlled steps.
// PASS 0: FP32 trace of every ray; 1: FP32 trace deferring photon-band rays;
// 2: FP64 trace of every ray; 3: FP64 trace of the deferred list.
template <int METRIC, int PASS>
__global__ __launch_bounds__(256) void trace_kernel(SceneConst<float> sf, SceneConst<double> sd,
                                                    const float2* __restrict__ halton, Chunk ch,
                                                    QueueCtl* __restrict__ ctl, std::uint32_t step_batch) {
    constexpr bool KERR = (METRIC >= 1);
    constexpr bool ADAPT = (METRIC == 2);
    constexpr bool F64 = (PASS >= 2);
    using S = typename std::conditional<F64, double, float>::type;
    const int lane = __lane_id();
    const std::uint32_t W = sf.width, H = sf.height, spp = sf.spp;
    const std::uint64_t n_items =
        (PASS == 3) ? std::uint64_t(ctl->band_count) : std::uint64_t(ch.S) * ch.n_tiles * 64u;
    std::uint64_t* counter = (PASS == 3) ? &ctl->next_band : &ctl->next_item;

    bool retired = false, live = false, need = true;
    std::uint32_t rec_idx = 0;
    std::uint64_t lane_steps = 0, lane_rays = 0;
    std::uint32_t lane_rej = 0;
    RayState<S> y{};

    for (;;) {
        // (1) claim rays for idle lanes with one atomic per wave
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
                        rec_idx = ch.band[item];
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
                        bool started;
                        bool deferred = false;
                        if (F64) {
                            started = init_ray<double, S, KERR>(y, sd, gx, gy, double(ox), double(oy), (double*)nullptr);
                        } else if (PASS == 1) {
                            float margin = 1e30f;
                            started = init_ray<float, S, KERR>(y, sf, gx, gy, ox, oy, &margin);
                            deferred = started && margin < sf.band;
                        } else {
                            started = init_ray<float, S, KERR>(y, sf, gx, gy, ox, oy, (float*)nullptr);
                        }
                        need = false;
                        if (deferred) {
                            const std::uint32_t slot = atomicAdd(&ctl->band_count, 1u);
                            ch.band[slot] = rec_idx;
                            need = true;
                        } else if (!started) {
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
        if (__ballot(!retired) == 0ull) break;

        // (2) integrate a batch of steps
        std::uint32_t term = kTermLive;
        for (std::uint32_t k = 0; k < step_batch; ++k) {
            if (live) {
                const std::uint32_t t = F64 ? advance<double, S, KERR, false, ADAPT>(y, sd, &lane_rej)
                                            : advance<float, S, KERR, PASS == 1, ADAPT>(y, sf, &lane_rej);
                if (t != kTermLive) {
                    live = false;
                    term = t;
                }
            }
        }

        // (3) emit finished rays; in the mixed pass, hand ambiguous outcomes to FP64
        if (term != kTermLive) {
            bool defer = false;
            if (PASS == 1) defer = float(y.near_eq) < kGrazeTheta;
            if (defer) {
                const std::uint32_t slot = atomicAdd(&ctl->band_count, 1u);
                ch.band[slot] = rec_idx;
            } else {
                ch.rec[rec_idx] = make_rec<S, KERR, ADAPT>(y, term);
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
    if (ADAPT && lane == 0 && rej) atomicAdd((unsigned long long*)&ctl->rejected, (unsigned long long)rej);
    if (lane == 0 && lane_rays) {
        atomicAdd((unsigned long long*)&ctl->steps, (unsigned long long)lane_steps);
        atomicAdd((unsigned long long*)&ctl->rays, (unsigned long long)lane_rays);
        if (F64) atomicAdd((unsigned long long*)&ctl->rays64, (unsigned long long)lane_rays);
    }
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

__device__ __forceinline__ std::uint8_t class_of(std::uint32_t term) {
    switch (term) {
        case kTermHorizon: return SING_CLASS_HORIZON;
        case kTermEscape: return SING_CLASS_ESCAPE;
        case kTermDisc: return SING_CLASS_DISC;
        case kTermBudget: return SING_CLASS_BUDGET;
        default: return SING_CLASS_INITFAIL;
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
                              hipStream_t stream) {
    const dim3 block(256);
    const std::uint32_t sb = shape.step_batch ? shape.step_batch : 8u;
    const bool kerr = sf.metric == SING_METRIC_KERR;
    const bool adapt = kerr && sf.tol > 0.0f;
#define SING_TRACE(P, G)                                                                                  \
    do {                                                                                                  \
        if (adapt)                                                                                        \
            hipLaunchKernelGGL((trace_kernel<2, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, sb); \
        else if (kerr)                                                                                    \
            hipLaunchKernelGGL((trace_kernel<1, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, sb); \
        else                                                                                              \
            hipLaunchKernelGGL((trace_kernel<0, P>), dim3(G), block, 0, stream, sf, sd, halton, ch, ctl, sb); \
    } while (0)
    if (mode == SING_PREC_FP32) {
        SING_TRACE(0, shape.blocks);
    } else if (mode == SING_PREC_FP64) {
        SING_TRACE(2, shape.blocks64);
    } else {
        SING_TRACE(1, shape.blocks);
        if (sf.flags & SING_FLAG_STARFIELD_ON)
            hipLaunchKernelGGL(audit_kernel, dim3(shape.blocks), block, 0, stream, sf, ch, ctl);
        SING_TRACE(3, shape.blocks64);
    }
#undef SING_TRACE
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
