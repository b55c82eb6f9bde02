// singularity_amd/csrc/trace.hip — the persistent wave64 trace kernels.
//
// One launch renders a frame.  The grid is persistent: each lane owns one
// pixel at a time and integrates that pixel's samples back to back, in sample
// order, so the per-pixel mean is accumulated in a fixed order (no float
// atomics) exactly like the reference's per-pixel sample loop
// (render/cuda/kernels/geodesic_kernel.cu:462-510).  When a lane finishes its
// pixel it takes the next one from a global queue; the wave claims new pixels
// for all its idle lanes with a single atomic, so long rays never hold the
// other 63 lanes of the wave hostage.  Rays are advanced in batches of
// `step_batch` RK4 steps between refills.
//
// PREC: 0 = FP32, 1 = FP64, 2 = mixed (FP32 bulk, FP64 for rays whose
// photon-orbit margin is below the band threshold).
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
    S v0, v1, v2, v3;     // Kerr: p_r, p_theta, E, L_z.  Schwarzschild: u^t, u^r, u^theta, u^phi
    S prev_r, prev_th, prev_uph;
    S aux0, aux1;         // disc crossing radius, u^phi at the crossing (Schwarzschild)
    std::uint32_t steps;
};

__device__ __forceinline__ std::uint64_t shfl_u64(std::uint64_t v, int src) {
    const std::uint32_t lo = __shfl(std::uint32_t(v), src);
    const std::uint32_t hi = __shfl(std::uint32_t(v >> 32), src);
    return (std::uint64_t(hi) << 32) | lo;
}

// Termination checks in the reference's order (geodesic_kernel.cu:351-401)
// followed by one RK4 step.  Arithmetic in R, storage in S.
template <class R, class S, bool KERR>
__device__ __forceinline__ std::uint32_t advance(RayState<S>& y, const SceneConst<R>& sc) {
    if (y.steps >= sc.max_steps) return kTermBudget;
    R r = R(y.r), th = R(y.th);
    if (!__builtin_isfinite(r) || r < sc.cut) return kTermHorizon;
    if (r > sc.escape_r) return kTermEscape;
    if (sc.flags & SING_FLAG_DISC_ON) {
        const R mid = R(1.5707963267948966192313216916397514);
        const R pth = R(y.prev_th);
        const R dp = pth - mid, dc = th - mid;
        if (dp * dc < R(0)) {
            const R frac = (mid - pth) / (th - pth);
            const R pr0 = R(y.prev_r);
            const R rc = pr0 + frac * (r - pr0);
            if (rc >= sc.r_in && rc <= sc.r_out) {
                y.aux0 = S(rc);
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
        kerr_rk4(r, th, ph, pr, pt, sc.h, sc, k);
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
    return kTermLive;
}

// Ray initialisation from the camera (geodesic_kernel.cu:229-242, 303-347).
// Returns false on a Kerr initialisation failure (no real root at the camera).
template <class R, class S, bool KERR>
__device__ __forceinline__ bool init_ray(RayState<S>& y, const SceneConst<R>& sc, std::uint32_t gx, std::uint32_t gy,
                                         R ox, R oy, double* margin_out) {
    R ur, uth, uph;
    camera_ray(sc, gx, gy, ox, oy, ur, uth, uph);
    y.r = S(sc.cam_r);
    y.th = S(sc.cam_theta);
    y.ph = S(sc.cam_phi);
    y.prev_r = y.r;
    y.prev_th = y.th;
    y.steps = 0;
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
            const double Ed = double(E), Ld = double(L), ptd = double(pt);
            const double Q = ptd * ptd + double(sc.q_cos2) * (Ld * Ld * double(sc.q_inv_s2) - double(sc.a2) * Ed * Ed);
            const double a = double(sc.a), M = double(sc.M);
            const double rp = M + sqrt(fmax(M * M - a * a, 0.0));
            *margin_out = photon_margin(Ld / Ed, Q / (Ed * Ed), a, M, rp);
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
            const double crd = double(cr), s = double(sc.sin_th);
            const double E = double(sc.f_cam) * double(ut);
            const double L2 = crd * crd * crd * crd * (double(uth) * double(uth) + s * s * double(uph) * double(uph));
            *margin_out = photon_margin(0.0, L2 / (E * E), 0.0, double(sc.M), 2.0 * double(sc.M));
        }
    }
    return true;
}

template <class R, class S, bool KERR>
__device__ __forceinline__ void shade(const RayState<S>& y, std::uint32_t term, const SceneConst<R>& sc, R* rgb) {
    rgb[0] = rgb[1] = rgb[2] = R(0);
    if (term == kTermEscape) {
        shade_escape(sc, R(y.th), R(y.ph), rgb);
    } else if (term == kTermDisc) {
        // Kerr: phi at the crossing is the current phi (geodesic_kernel.cu:389);
        // Schwarzschild: lerp(phi, phi) == phi (geodesic_kernel.cu:282).
        shade_disc<R, KERR>(sc, R(y.aux0), R(y.ph), KERR ? R(y.v3) : R(y.aux1), rgb);
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

template <int METRIC, int PREC>
__global__ __launch_bounds__(256) void trace_kernel(SceneConst<float> sf, SceneConst<double> sd,
                                                    const float2* __restrict__ halton, TraceOutputs out,
                                                    QueueCtl* __restrict__ ctl, std::uint32_t n_items,
                                                    std::uint32_t tiles_x, std::uint32_t step_batch) {
    constexpr bool KERR = (METRIC == 1);
    using S = typename std::conditional<PREC == 0, float, double>::type;   // ray storage
    using A = S;                                                            // accumulators
    const int lane = __lane_id();
    const std::uint32_t W = sf.width, H = sf.height, spp = sf.spp;

    bool retired = false, need_pixel = true, live = false, use64 = (PREC == 1);
    std::uint32_t term = kTermLive, gx = 0, gy = 0, s = 0, px_steps = 0;
    std::uint8_t cls0 = 0;
    A ldr0 = A(0), ldr1 = A(0), ldr2 = A(0), hdr0 = A(0), hdr1 = A(0), hdr2 = A(0);
    std::uint64_t lane_steps = 0, lane_rays = 0, lane_rays64 = 0;
    RayState<S> y{};

    for (;;) {
        // (1) claim pixels for idle lanes with one atomic per wave
        const std::uint64_t need = __ballot(need_pixel && !retired);
        if (need) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            std::uint64_t base = 0;
            if (lane == leader) base = atomicAdd((unsigned long long*)&ctl->next_item, (unsigned long long)__popcll(need));
            base = shfl_u64(base, leader);
            if (need_pixel && !retired) {
                const std::uint64_t item = base + __popcll(need & ((1ull << lane) - 1ull));
                if (item >= n_items) {
                    retired = true;
                } else {
                    const std::uint32_t tile = std::uint32_t(item >> 6), w = std::uint32_t(item & 63u);
                    gx = (tile % tiles_x) * 8u + (w & 7u);
                    gy = (tile / tiles_x) * 8u + (w >> 3);
                    if (gx < W && gy < H) {
                        need_pixel = false;
                        s = 0;
                        px_steps = 0;
                        ldr0 = ldr1 = ldr2 = hdr0 = hdr1 = hdr2 = A(0);
                    }
                }
            }
        }
        if (__ballot(!retired) == 0ull) break;

        // (2) start the next sample of the lane's pixel
        if (!retired && !need_pixel && !live) {
            float ox = 0.5f, oy = 0.5f;
            if (spp > 1u) {
                const float2 hv = halton[s];
                ox = hv.x;
                oy = hv.y;
            }
            bool ok;
            if (PREC == 0) {
                ok = init_ray<float, S, KERR>(y, sf, gx, gy, ox, oy, nullptr);
            } else if (PREC == 1) {
                ok = init_ray<double, S, KERR>(y, sd, gx, gy, double(ox), double(oy), nullptr);
            } else {
                double margin = 1e30;
                ok = init_ray<double, S, KERR>(y, sd, gx, gy, double(ox), double(oy), &margin);
                use64 = margin < sd.band;
            }
            live = ok;
            term = ok ? kTermLive : kTermInitFail;
        }

        // (3) integrate a batch of steps
        for (std::uint32_t k = 0; k < step_batch; ++k) {
            if (live) {
                std::uint32_t t;
                if (PREC == 0) t = advance<float, S, KERR>(y, sf);
                else if (PREC == 1) t = advance<double, S, KERR>(y, sd);
                else t = use64 ? advance<double, S, KERR>(y, sd) : advance<float, S, KERR>(y, sf);
                if (t != kTermLive) {
                    live = false;
                    term = t;
                }
            }
        }

        // (4) shade and accumulate finished rays, emit finished pixels
        if (!retired && !need_pixel && !live) {
            A rgb[3];
            if (PREC == 0) {
                float c[3];
                shade<float, S, KERR>(y, term, sf, c);
                rgb[0] = c[0]; rgb[1] = c[1]; rgb[2] = c[2];
            } else {
                double c[3];
                shade<double, S, KERR>(y, term, sd, c);
                rgb[0] = A(c[0]); rgb[1] = A(c[1]); rgb[2] = A(c[2]);
            }
            const A ex = (PREC == 0) ? A(sf.exposure) : A(sd.exposure);
            hdr0 += rgb[0];
            hdr1 += rgb[1];
            hdr2 += rgb[2];
            ldr0 += aces1(rgb[0] * ex);
            ldr1 += aces1(rgb[1] * ex);
            ldr2 += aces1(rgb[2] * ex);
            if (s == 0) cls0 = class_of(term);
            px_steps += y.steps;
            lane_steps += y.steps;
            lane_rays += 1;
            if (PREC == 2 && use64) lane_rays64 += 1;
            if (PREC == 1) lane_rays64 += 1;
            term = kTermLive;
            s += 1;
            if (s == spp) {
                const std::uint32_t pix = gy * W + gx;
                const A inv = (spp > 1u) ? A(1) / A(spp) : A(1);
                const A l0 = ldr0 * inv, l1 = ldr1 * inv, l2 = ldr2 * inv;
                if (out.rgba) out.rgba[pix] = make_uchar4(to_u8(l0), to_u8(l1), to_u8(l2), 255);
                const float h0 = float(hdr0 * inv), h1 = float(hdr1 * inv), h2 = float(hdr2 * inv);
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
                if (out.cls) out.cls[pix] = cls0;
                if (out.steps) out.steps[pix] = px_steps;
                need_pixel = true;
            }
        }
    }

    // per-wave totals
    for (int off = 32; off > 0; off >>= 1) {
        lane_steps += shfl_u64(lane_steps, lane ^ off);
        lane_rays += shfl_u64(lane_rays, lane ^ off);
        lane_rays64 += shfl_u64(lane_rays64, lane ^ off);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long*)&ctl->steps, (unsigned long long)lane_steps);
        atomicAdd((unsigned long long*)&ctl->rays, (unsigned long long)lane_rays);
        atomicAdd((unsigned long long*)&ctl->rays64, (unsigned long long)lane_rays64);
    }
}

}  // namespace

hipError_t launch_trace(const SceneConst<float>& sf, const SceneConst<double>& sd, const float2* halton,
                        const TraceOutputs& out, QueueCtl* ctl, const LaunchShape& shape, hipStream_t stream) {
    const std::uint32_t tiles_x = (sf.width + 7u) / 8u, tiles_y = (sf.height + 7u) / 8u;
    const std::uint32_t n_items = tiles_x * tiles_y * 64u;
    const dim3 block(256), grid(shape.blocks);
    const std::uint32_t sb = shape.step_batch ? shape.step_batch : 8u;
    const bool kerr = sf.metric == SING_METRIC_KERR;
#define SING_LAUNCH(M, P) \
    hipLaunchKernelGGL((trace_kernel<M, P>), grid, block, 0, stream, sf, sd, halton, out, ctl, n_items, tiles_x, sb)
    switch (shape.precision) {
        case SING_PREC_FP32: if (kerr) SING_LAUNCH(1, 0); else SING_LAUNCH(0, 0); break;
        case SING_PREC_FP64: if (kerr) SING_LAUNCH(1, 1); else SING_LAUNCH(0, 1); break;
        default: if (kerr) SING_LAUNCH(1, 2); else SING_LAUNCH(0, 2); break;
    }
#undef SING_LAUNCH
    return hipGetLastError();
}

}  // namespace sing
