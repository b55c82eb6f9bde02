// singularity_amd/csrc/adjoint.hip — differentiable render: forward pass with
// checkpoints and a hand-written discrete adjoint of the RK4 march.
//
// Forward (one thread per pixel, 1 sample at the pixel centre): initialise the
// ray, march with the damped Kerr RK4 (kerr_hamilton.h:165-178), store the
// state every `C` steps, and record every equator crossing that meets the
// soft annulus (step index, opacity, emission, crossing radius).  Radiance is
// the front-to-back composite of the crossings (plus the escape colour behind
// them); the loss is sum_pixels w_p |I_p - T_p|^2.
//
// Backward (one thread per pixel): the crossing gradients follow from the
// composite; emission and opacity Jacobians come from forward-mode duals; the
// state adjoint is swept back through the march segment by segment, each
// segment recomputed from its checkpoint and reversed with kerr_step_vjp.  The
// camera initialisation's dependence on (a, inclination) is differentiated in
// forward mode.  Per-pixel gradients are reduced in a fixed order.
//
// Mixed precision: the FP32 pass hands captured, photon-band and grazing rays
// to an FP64 pass, exactly as the render path does; each ray's backward runs in
// the precision of its forward.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../../include/sing_abi.h"
#include "adjoint.hpp"
#include "context.hpp"
#include "diff_scene.hpp"
#include "physics.hpp"

namespace sing {

namespace {

// Mixed precision (as the render): rays that hard_ray flags (photon-orbit
// margin below kBandDiff, reaching 2 r+, reaching 0.03 rad of the polar axis)
// go to FP64 before the march; after it, a state within kGrazeThetaDiff of the
// equator or 0.03 rad of the polar axis, or a ray with emission that came
// within 2 r+ of the hole (both caught before the march unless rounding moves
// the ray across the threshold).
constexpr float kGrazeThetaDiff = 2e-6f;
constexpr float kBandDiff = 0.03f;
constexpr float kPoleGapDiff = 1.5707963267948966f - 0.03f;
constexpr std::uint32_t kMetaPrec64 = 1u;  // meta[3]: precision of the pixel's forward

template <class R> __device__ __forceinline__ const DiffScene<R>& scene_of(const DiffSet* ds);
template <> __device__ __forceinline__ const DiffScene<float>& scene_of<float>(const DiffSet* ds) { return ds->sf; }
template <> __device__ __forceinline__ const DiffScene<double>& scene_of<double>(const DiffSet* ds) { return ds->sd; }
template <class R> __device__ __forceinline__ const DiffScene<Dual<2, R>>& tangent_scene_of(const DiffSet* ds);
template <> __device__ __forceinline__ const DiffScene<Dual<2, float>>& tangent_scene_of<float>(const DiffSet* ds) {
    return ds->tf;
}
template <> __device__ __forceinline__ const DiffScene<Dual<2, double>>& tangent_scene_of<double>(const DiffSet* ds) {
    return ds->td;
}

// Per-pixel arrays are slot-major with the pixel index fastest (element k of
// slot s of pixel p at [(s * width + k) * P + p]), so a wavefront's accesses to
// one element are contiguous.
template <class R> __device__ __forceinline__ void store_state(double* p, std::size_t S, const KY<R>& y) {
    p[0] = double(y.r);
    p[S] = double(y.th);
    p[2 * S] = double(y.ph);
    p[3 * S] = double(y.pr);
    p[4 * S] = double(y.pt);
}
template <class R> __device__ __forceinline__ KY<R> load_state(const double* p, std::size_t S) {
    return {R(p[0]), R(p[S]), R(p[2 * S]), R(p[3 * S]), R(p[4 * S])};
}

// Mixed precision, first: rays that hard_ray flags go to the FP64 list (and
// meta[3] = FP64) before any march, so the FP32 and FP64 forwards can run
// concurrently.
__global__ __launch_bounds__(256) void diff_classify_kernel(const DiffSet* __restrict__ ds, DiffBuffers b) {
    const std::uint32_t ev = blockIdx.y, pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= b.Pimg) return;
    const std::uint32_t q = ev * b.Pimg + pix;
    const DiffScene<float>& sc = ds[ev].sf;
    KY<float> y;
    float E, L;
    const bool ok = diff_init(sc, pix % sc.width, pix / sc.width, y, E, L);
    bool hard = false;
    if (ok) {
        float st, ct;
        dsincos(sc.cam_theta, &st, &ct);
        const float Q = y.pt * y.pt + ct * ct * (L * L / fmaxf(st * st, 1e-6f) - sc.a * sc.a * E * E);
        const float rp = sc.M + sqrtf(fmaxf(sc.M * sc.M - sc.a * sc.a, 0.0f));
        hard = hard_ray<float>(L / E, Q / (E * E), sc.a, sc.M, rp, y.pr < 0.0f, kBandDiff);
    }
    b.meta[4u * q + 3u] = hard ? kMetaPrec64 : 0u;
    if (hard) b.list[ev * b.Pimg + atomicAdd(b.count + ev, 1u)] = pix;
}

// MODE 0: FP32 over all pixels; 1: FP32 over the pixels the classification
// left in FP32, handing grazing rays to the late list (list2); 2: FP64 over all
// pixels; 3: FP64 over a list (b.list / b.count).  Grid: (pixels of one image,
// evaluations); blockIdx.y is the evaluation, so its scene stays wave-uniform.
template <class R, int MODE>
__global__ __launch_bounds__(256) void diff_fwd_kernel(const DiffSet* __restrict__ ds, DiffBuffers b) {
    // the mixed form's FP64 lists (MODE 3) hold its longest rays, which set the
    // evaluation's latency while FP32 waves share their SIMDs: issue first
    if (MODE == 3) __builtin_amdgcn_s_setprio(2);
    const std::uint32_t ev = blockIdx.y, idx = blockIdx.x * 256u + threadIdx.x;
    std::uint32_t pix;
    if (MODE == 3) {
        if (idx >= b.count[ev]) return;
        pix = b.list[ev * b.Pimg + idx];
    } else {
        if (idx >= b.Pimg) return;
        pix = idx;
    }
    const std::uint32_t q = ev * b.Pimg + pix;  // index into the per-pixel arrays
    const DiffScene<R>& sc = scene_of<R>(ds + ev);
    const std::uint32_t gx = pix % sc.width, gy = pix / sc.width;
    const std::size_t S = b.P;
    std::uint32_t* meta = b.meta + 4u * q;
    if (MODE == 1 && meta[3] == kMetaPrec64) return;  // classified hard: the FP64 forward owns it
    double* cp = b.cp + q;
    double* cr = b.cross + q;
    KY<R> y;
    R E, L;
    const bool ok = diff_init(sc, gx, gy, y, E, L);
    if (MODE != 1) meta[3] = (MODE >= 2) ? kMetaPrec64 : 0u;
    R acc[3] = {R(0), R(0), R(0)};
    R tau = R(1);
    std::uint32_t ncross = 0, step = 0, term = kTermInitFail;
    double sky[3] = {0.0, 0.0, 0.0};
    R near_eq = R(1e30), rmin = R(1e30);
    if (ok) {
        const KC<R> c{E, L, sc.a, sc.M};
        b.EL[q] = double(E);
        b.EL[S + q] = double(L);
        store_state(cp, S, y);
        R prev_r = y.r, prev_th = y.th, prev_ph = y.ph;
        const R mid = R(1.5707963267948966192313216916397514);
        for (;;) {
            if (step >= sc.max_steps) {
                term = kTermBudget;
                break;
            }
            if (!__builtin_isfinite(y.r) || y.r < sc.cut) {
                term = kTermHorizon;
                break;
            }
            if (y.r > sc.escape_r) {
                term = kTermEscape;
                break;
            }
            if (sc.flags & SING_FLAG_DISC_ON) {
                const R dp = prev_th - mid, dc = y.th - mid;
                if (dp * dc < R(0)) {
                    const R frac = (mid - prev_th) / (y.th - prev_th);
                    const R rc = prev_r + frac * (y.r - prev_r);
                    const R alpha = soft_alpha(rc, sc.r_in, sc.r_out, sc.w);
                    if (alpha > R(0)) {
                        const R phc = prev_ph + frac * (y.ph - prev_ph);
                        R col[3];
                        diff_emit(rc, phc, L, sc.a, sc.M, sc.r_in, sc.peak_T, sc.f_obs, sc.turbulence, sc.time_sec,
                                  sc.flags, col);
                        double* rec = cr + 7u * ncross * S;
                        rec[0] = double(step);
                        rec[S] = double(alpha);
                        rec[2 * S] = double(col[0]);
                        rec[3 * S] = double(col[1]);
                        rec[4 * S] = double(col[2]);
                        rec[5 * S] = double(rc);
                        rec[6 * S] = double(phc);
                        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + tau * alpha * col[k];
                        ++ncross;
                        if (!(alpha < R(1)) || ncross == std::uint32_t(kMaxCross)) {
                            term = kTermDisc;
                            break;
                        }
                        tau = tau * (R(1) - alpha);
                    }
                }
            }
            prev_th = y.th;
            prev_r = y.r;
            prev_ph = y.ph;
            y = kerr_step(y, c, sc.h);
            ++step;
            if (step % b.C == 0u) store_state(cp + 5u * (step / b.C) * S, S, y);
            if (MODE == 1) {
                const R d = y.th - mid;
                const R ad = d < R(0) ? -d : d;
                if (ad < near_eq) near_eq = ad;
                if (ad > R(kPoleGapDiff)) near_eq = R(0);  // near the polar axis
                if (y.r < rmin) rmin = y.r;
            }
        }
        const R rp = sc.M + hsqrt(hmax(sc.M * sc.M - sc.a * sc.a, R(0)));
        if (MODE == 1 && (float(near_eq) < kGrazeThetaDiff || (ncross > 0u && rmin < R(2) * rp))) {
            meta[3] = kMetaPrec64;
            b.list2[ev * b.Pimg + atomicAdd(b.count2 + ev, 1u)] = pix;
            return;
        }
        if (term == kTermEscape && (sc.flags & SING_FLAG_STARFIELD_ON)) {
            SceneConst<float> sk{};
            sk.flags = sc.flags;
            float c3[3];
            shade_escape<float>(sk, float(y.th), float(y.ph), c3);
            for (int k = 0; k < 3; ++k) {
                sky[k] = c3[k];
                acc[k] = acc[k] + tau * R(c3[k]);
            }
        }
    }
    meta[0] = step;
    meta[1] = ncross;
    meta[2] = term;
    double* skp = b.sky + 3u * q;
    double l = 0.0;
    if (b.cot) {
        const float* g = b.cot + 3u * pix;
        for (int k = 0; k < 3; ++k) {
            skp[k] = sky[k];
            b.hdr[3u * q + k] = float(acc[k]);
            b.gI[3u * q + k] = double(g[k]);
            l += double(g[k]) * double(acc[k]);
        }
    } else {
        const std::size_t tp = std::size_t(ev) * b.tstride + pix;
        const float* tgt = b.target + 3u * tp;
        const double w = b.weight ? double(b.weight[tp]) : 1.0;
        for (int k = 0; k < 3; ++k) {
            skp[k] = sky[k];
            b.hdr[3u * q + k] = float(acc[k]);
            const double diff = double(acc[k]) - double(tgt[k]);
            b.gI[3u * q + k] = 2.0 * w * diff;
            l += w * diff * diff;
        }
    }
    b.loss[q] = l;
    b.steps[q] = step;
}

// MODE 0: every pixel whose forward ran in R's precision; 1: the pixels of
// b.list (the mixed form's FP64 lists).
template <class R, int MODE>
__global__ __launch_bounds__(256) void diff_bwd_kernel(const DiffSet* __restrict__ ds, DiffBuffers b) {
    if (MODE == 1) __builtin_amdgcn_s_setprio(2);  // the FP64 lists, as in the forward
    const std::uint32_t ev = blockIdx.y, idx = blockIdx.x * 256u + threadIdx.x;
    std::uint32_t pix;
    if (MODE == 1) {
        if (idx >= b.count[ev]) return;
        pix = b.list[ev * b.Pimg + idx];
    } else {
        if (idx >= b.Pimg) return;
        pix = idx;
    }
    const std::uint32_t q = ev * b.Pimg + pix;  // index into the per-pixel arrays
    ds += ev;
    const std::uint32_t* meta = b.meta + 4u * q;
    double* g = b.grad + 5u * q;
    const bool want64 = std::is_same<R, double>::value;
    if ((meta[3] == kMetaPrec64) != want64) return;
    for (int k = 0; k < 5; ++k) g[k] = 0.0;
    const DiffScene<R>& sc = scene_of<R>(ds);
    const std::uint32_t gx = pix % sc.width, gy = pix / sc.width;
    // camera initialisation and its tangents wrt (a/M, inclination)
    KY<Dual<2, R>> y0;
    Dual<2, R> Ed, Ld;
    if (!diff_init(tangent_scene_of<R>(ds), gx, gy, y0, Ed, Ld)) return;
    const std::uint32_t J = meta[1], term = meta[2];
    if (J == 0u) return;  // no crossing: the radiance has no parameter dependence
    const std::size_t S = b.P;
    // the forward's own constants, so the recomputed segments repeat it exactly
    const KC<R> c{R(b.EL[q]), R(b.EL[S + q]), sc.a, sc.M};
    const double* gI = b.gI + 3u * q;
    const double* cr = b.cross + q;

    // ---- crossing gradients (front-to-back composite)
    R tau[kMaxCross];
    R t = R(1);
    for (std::uint32_t j = 0; j < J; ++j) {
        tau[j] = t;
        t = t * (R(1) - R(cr[(7u * j + 1u) * S]));
    }
    const bool sky_behind = (term == kTermEscape);
    R behind[3] = {R(0), R(0), R(0)};
    if (sky_behind)
        for (int k = 0; k < 3; ++k) behind[k] = R(b.sky[3u * q + k]);
    R rc_bar[kMaxCross], ph_bar[kMaxCross];
    R gL = R(0), g_a = R(0), g_rin = R(0), g_rout = R(0), g_ts = R(0);
    const R Ts = R(ds->tscale), T0 = R(ds->T0);
    for (int j = int(J) - 1; j >= 0; --j) {
        const double* rec = cr + 7u * std::size_t(j) * S;
        const R alpha = R(rec[S]);
        const R col[3] = {R(rec[2 * S]), R(rec[3 * S]), R(rec[4 * S])};
        R cbar[3], abar = R(0);
        for (int k = 0; k < 3; ++k) {
            cbar[k] = R(gI[k]) * tau[j] * alpha;
            abar = abar + R(gI[k]) * tau[j] * (col[k] - behind[k]);
        }
        for (int k = 0; k < 3; ++k) behind[k] = alpha * col[k] + (R(1) - alpha) * behind[k];
        // emission Jacobian wrt (rc, phi, L, a, r_in, Ts) at the recorded crossing
        using D6 = Dual<6, R>;
        const R rc = R(rec[5 * S]);
        const R phi_k = R(rec[6 * S]);
        D6 rgb[3];
        diff_emit<D6>(D6::var(rc, 0), D6::var(phi_k, 1), D6::var(c.L, 2), D6::var(sc.a, 3), D6(sc.M),
                      D6::var(sc.r_in, 4), D6(T0) * D6::var(Ts, 5), D6(sc.f_obs), D6(sc.turbulence), D6(sc.time_sec),
                      sc.flags, rgb);
        using D3 = Dual<3, R>;
        const D3 al = soft_alpha(D3::var(rc, 0), D3::var(sc.r_in, 1), D3::var(sc.r_out, 2), D3(sc.w));
        R grc = abar * al.d[0], gph = R(0);
        for (int k = 0; k < 3; ++k) {
            grc = grc + cbar[k] * rgb[k].d[0];
            gph = gph + cbar[k] * rgb[k].d[1];
            gL = gL + cbar[k] * rgb[k].d[2];
            g_a = g_a + cbar[k] * rgb[k].d[3];
            g_rin = g_rin + cbar[k] * rgb[k].d[4];
            g_ts = g_ts + cbar[k] * rgb[k].d[5];
        }
        g_rin = g_rin + abar * al.d[1];
        g_rout = g_rout + abar * al.d[2];
        rc_bar[j] = grc;
        ph_bar[j] = gph;
    }

    // ---- adjoint sweep through the march
    KY<R> lam = {R(0), R(0), R(0), R(0), R(0)};
    KG<R> gc = {R(0), R(0), R(0)};
    gc.L = gL;
    const double* cp = b.cp + q;
    double* scr = b.scratch + q;
    int jn = int(J) - 1;  // next crossing (from the end) whose seeds are pending
    // The state adjoint is zero after the last crossing, so the sweep starts
    // there.  (Sweeping the remaining steps would only add 0 * J^T terms, and
    // the final step into the horizon can have non-finite Jacobian entries.)
    const std::uint32_t n = std::uint32_t(cr[7u * std::size_t(jn) * S]);
    if (n > 0) {
        const std::uint32_t nseg = (n + b.C - 1u) / b.C;
        for (int seg = int(nseg) - 1; seg >= 0; --seg) {
            const std::uint32_t k0 = std::uint32_t(seg) * b.C;
            const std::uint32_t k1 = (k0 + b.C < n) ? k0 + b.C : n;
            KY<R> ys = load_state<R>(cp + 5u * std::size_t(seg) * S, S);
            store_state(scr, S, ys);
            for (std::uint32_t i = 1; i <= k1 - k0; ++i) {
                ys = kerr_step(ys, c, sc.h);
                store_state(scr + 5u * i * S, S, ys);
            }
            for (std::uint32_t k = k1; k > k0; --k) {
                const KY<R> yk = load_state<R>(scr + 5u * (k - k0) * S, S);
                const KY<R> ykm = load_state<R>(scr + 5u * (k - 1u - k0) * S, S);
                // crossing between y_{k-1} and y_k: x_c = x_{k-1} + f (x_k - x_{k-1}) for
                // x = r, phi with f = (pi/2 - theta_{k-1}) / (theta_k - theta_{k-1})
                R seed_prev_r = R(0), seed_prev_th = R(0), seed_prev_ph = R(0);
                bool has = false;
                if (jn >= 0 && std::uint32_t(cr[7u * std::size_t(jn) * S]) == k) {
                    has = true;
                    const R mid = R(1.5707963267948966192313216916397514);
                    const R dth = yk.th - ykm.th;
                    const R f = (mid - ykm.th) / dth;
                    const R rb = rc_bar[jn], pb = ph_bar[jn];
                    const R dx = rb * (yk.r - ykm.r) + pb * (yk.ph - ykm.ph);  // x_bar . (x_k - x_{k-1})
                    lam.r = lam.r + rb * f;
                    lam.ph = lam.ph + pb * f;
                    lam.th = lam.th + dx * (-f / dth);
                    seed_prev_r = rb * (R(1) - f);
                    seed_prev_ph = pb * (R(1) - f);
                    seed_prev_th = dx * (f - R(1)) / dth;
                    --jn;
                }
                lam = kerr_step_vjp(ykm, c, sc.h, lam, gc);
                if (has) {
                    lam.r = lam.r + seed_prev_r;
                    lam.th = lam.th + seed_prev_th;
                    lam.ph = lam.ph + seed_prev_ph;
                }
            }
        }
    }

    // ---- camera initialisation: dJ/dp = lam . dy0/dp + gE dE/dp + gL dL/dp (+ direct terms)
    const R M = sc.M;
    for (int p = 0; p < 2; ++p) {
        R v = lam.r * y0.r.d[p] + lam.th * y0.th.d[p] + lam.ph * y0.ph.d[p] + lam.pr * y0.pr.d[p] +
              lam.pt * y0.pt.d[p] + gc.E * Ed.d[p] + gc.L * Ld.d[p];
        g[p] = double(v);
    }
    g[0] += double((gc.a + g_a) * M);  // a = (a/M) M
    g[2] = double(g_rin * M);
    g[3] = double(g_rout * M);
    g[4] = double(g_ts);
}

// Loss and gradient sums in a fixed order (bit-reproducible, independent of
// scheduling): stage 1, kReduceBlocks blocks each sum a contiguous slice of
// pixels (a block-stride loop, then a tree in LDS); stage 2, one block sums
// the per-block partials in block order.
// The step total is a sum of integers below 2^53, exact in double.
// Block (x, e) sums slice x of evaluation e's pixels (Pimg per evaluation).
__global__ __launch_bounds__(256) void diff_reduce_partial_kernel(const double* __restrict__ loss,
                                                                  const double* __restrict__ grad,
                                                                  const std::uint32_t* __restrict__ steps,
                                                                  std::uint32_t Pimg, double* __restrict__ partial) {
    constexpr int V = int(kReduceVals);
    __shared__ double sh[V][256];
    const std::uint32_t t = threadIdx.x;
    const std::uint32_t per = (Pimg + kReduceBlocks - 1u) / kReduceBlocks;
    const std::uint32_t base = blockIdx.y * Pimg;
    const std::uint32_t lo = base + blockIdx.x * per, hi = base + min(Pimg, blockIdx.x * per + per);
    double s[V] = {0, 0, 0, 0, 0, 0, 0};
    for (std::uint32_t i = lo + t; i < hi; i += 256u) {
        s[0] += loss[i];
        for (int k = 0; k < 5; ++k) s[1 + k] += grad[5u * i + k];
        s[6] += double(steps[i]);
    }
    for (int k = 0; k < V; ++k) sh[k][t] = s[k];
    __syncthreads();
    for (std::uint32_t w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < V; ++k) sh[k][t] += sh[k][t + w];
        __syncthreads();
    }
    if (t < kReduceVals) partial[kReduceVals * (blockIdx.y * kReduceBlocks + blockIdx.x) + t] = sh[t][0];
}

// Block e: out[e * kDiffOut + 0..kReduceVals): evaluation e's sums;
// out[e * kDiffOut + kReduceVals]: the length of its FP64 list (0 without a
// list), so one copy brings back every per-evaluation result.
__global__ __launch_bounds__(256) void diff_reduce_final_kernel(const double* __restrict__ partial,
                                                                const std::uint32_t* __restrict__ count,
                                                                double* __restrict__ out) {
    constexpr int V = int(kReduceVals);
    __shared__ double sh[V][kReduceBlocks];
    const std::uint32_t t = threadIdx.x;
    partial += std::size_t(blockIdx.x) * kReduceBlocks * kReduceVals;
    out += std::size_t(blockIdx.x) * kDiffOut;
    for (int k = 0; k < V; ++k) sh[k][t] = partial[kReduceVals * t + k];
    __syncthreads();
    for (std::uint32_t w = kReduceBlocks / 2u; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < V; ++k) sh[k][t] += sh[k][t + w];
        __syncthreads();
    }
    if (t < kReduceVals) out[t] = sh[t][0];
    if (t == kReduceVals) out[t] = count ? double(count[blockIdx.x]) : 0.0;
}

}  // namespace

hipError_t launch_diff(const DiffSet* ds_dev, const DiffBuffers& b, std::uint32_t precision, bool backward,
                       double* out, hipStream_t stream, hipStream_t stream2, hipEvent_t fwd_end,
                       const hipEvent_t* join) {
    // fwd_end: timing event at the end of the forward (on `stream`); join[0..3]:
    // cross-stream joins, created without timing (they only order work)
    const dim3 block(256), grid((b.Pimg + 255u) / 256u, b.n_eval);
    if (precision == SING_PREC_FP32) {
        hipLaunchKernelGGL((diff_fwd_kernel<float, 0>), grid, block, 0, stream, ds_dev, b);
        (void)hipEventRecord(fwd_end, stream);
        if (backward) hipLaunchKernelGGL((diff_bwd_kernel<float, 0>), grid, block, 0, stream, ds_dev, b);
    } else if (precision == SING_PREC_FP64) {
        hipLaunchKernelGGL((diff_fwd_kernel<double, 2>), grid, block, 0, stream, ds_dev, b);
        (void)hipEventRecord(fwd_end, stream);
        if (backward) hipLaunchKernelGGL((diff_bwd_kernel<double, 0>), grid, block, 0, stream, ds_dev, b);
    } else {
        // Mixed: the hard rays (classified first) run in FP64 on stream2 while the
        // rest run in FP32 on stream; grazing rays found by the FP32 march follow
        // in FP64 (late list).  Each launch is latency-bound at these sizes (one
        // ray per thread), so running the two precisions side by side hides the
        // FP64 rays' long tail instead of adding it.
        DiffBuffers late = b;
        late.list = b.list2;
        late.count = b.count2;
        // A pixel's backward needs only its own forward (the loss gradient
        // dL/dI of the pixel is written at the end of its forward), so the
        // classified FP64 rays — the evaluation's longest — go straight on to
        // their backward on stream2 while stream runs the FP32 forward, the
        // late FP64 forward and both remaining backwards.
        hipLaunchKernelGGL(diff_classify_kernel, grid, block, 0, stream, ds_dev, b);
        (void)hipEventRecord(join[0], stream);
        (void)hipStreamWaitEvent(stream2, join[0], 0);
        hipLaunchKernelGGL((diff_fwd_kernel<double, 3>), grid, block, 0, stream2, ds_dev, b);
        (void)hipEventRecord(join[1], stream2);
        if (backward) hipLaunchKernelGGL((diff_bwd_kernel<double, 1>), grid, block, 0, stream2, ds_dev, b);
        (void)hipEventRecord(join[2], stream2);
        hipLaunchKernelGGL((diff_fwd_kernel<float, 1>), grid, block, 0, stream, ds_dev, b);
        hipLaunchKernelGGL((diff_fwd_kernel<double, 3>), grid, block, 0, stream, ds_dev, late);
        (void)hipStreamWaitEvent(stream, join[1], 0);
        (void)hipEventRecord(fwd_end, stream);
        if (backward) {
            hipLaunchKernelGGL((diff_bwd_kernel<float, 0>), grid, block, 0, stream, ds_dev, b);
            hipLaunchKernelGGL((diff_bwd_kernel<double, 1>), grid, block, 0, stream, ds_dev, late);
        }
        (void)hipStreamWaitEvent(stream, join[2], 0);
    }
    hipLaunchKernelGGL(diff_reduce_partial_kernel, dim3(kReduceBlocks, b.n_eval), dim3(256), 0, stream, b.loss,
                       b.grad, b.steps, b.Pimg, b.partial);
    hipLaunchKernelGGL(diff_reduce_final_kernel, dim3(b.n_eval), dim3(kReduceBlocks), 0, stream, b.partial,
                       precision == SING_PREC_MIXED ? b.count : nullptr, out);
    return hipGetLastError();
}

}  // namespace sing

// ============================================================= host side
void SingContext::free_diff() {
    (void)hipFree(d_dset);
    (void)hipFree(diff.cp);
    (void)hipFree(diff.cross);
    (void)hipFree(diff.meta);
    (void)hipFree(diff.sky);
    (void)hipFree(diff.EL);
    (void)hipFree(diff.gI);
    (void)hipFree(diff.hdr);
    (void)hipFree(diff.loss);
    (void)hipFree(diff.grad);
    (void)hipFree(diff.scratch);
    (void)hipFree(diff.list);
    (void)hipFree(diff.count);
    (void)hipFree(diff.list2);
    (void)hipFree(diff.count2);
    (void)hipFree(diff.steps);
    (void)hipFree(diff.partial);
    (void)hipFree(d_target);
    (void)hipFree(d_weight);
    (void)hipFree(d_out);
    if (h_out) (void)hipHostFree(h_out);
    d_dset = nullptr;
    diff = sing::DiffBuffers{};
    d_target = d_weight = nullptr;
    d_out = nullptr;
    h_out = nullptr;
    diff_P = diff_n = diff_ncp = diff_C = 0;
}

int SingContext::ensure_diff(std::size_t Pimg, std::size_t n_eval, std::size_t ncp, std::size_t C) {
    // buffers sized for n evaluations serve any smaller call (the per-pixel
    // arrays are indexed with the call's own P = n_eval * Pimg)
    if (Pimg == diff_P && n_eval <= diff_n && ncp <= diff_ncp && C == diff_C && d_dset) {
        diff.P = std::uint32_t(Pimg * n_eval);
        diff.n_eval = std::uint32_t(n_eval);
        return SING_OK;
    }
    free_diff();
    const std::size_t P = Pimg * n_eval;
    hipError_t e = hipSuccess;
    auto A = [&](void** p, std::size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    A((void**)&d_dset, n_eval * sizeof(sing::DiffSet));
    A((void**)&diff.cp, P * ncp * 5 * sizeof(double));
    A((void**)&diff.cross, P * sing::kMaxCross * 7 * sizeof(double));
    A((void**)&diff.meta, P * 4 * sizeof(std::uint32_t));
    A((void**)&diff.sky, P * 3 * sizeof(double));
    A((void**)&diff.EL, P * 2 * sizeof(double));
    A((void**)&diff.gI, P * 3 * sizeof(double));
    A((void**)&diff.hdr, P * 3 * sizeof(float));
    A((void**)&diff.loss, P * sizeof(double));
    A((void**)&diff.grad, P * 5 * sizeof(double));
    A((void**)&diff.scratch, P * (C + 1) * 5 * sizeof(double));
    A((void**)&diff.list, P * sizeof(std::uint32_t));
    A((void**)&diff.count, n_eval * sizeof(std::uint32_t));
    A((void**)&diff.list2, P * sizeof(std::uint32_t));
    A((void**)&diff.count2, n_eval * sizeof(std::uint32_t));
    A((void**)&diff.steps, P * sizeof(std::uint32_t));
    A((void**)&diff.partial, n_eval * sing::kReduceBlocks * sing::kReduceVals * sizeof(double));
    A((void**)&d_target, P * 3 * sizeof(float));  // one image per evaluation at most
    A((void**)&d_weight, P * sizeof(float));
    A((void**)&d_out, n_eval * sing::kDiffOut * sizeof(double));
    // pinned, so the per-call result copy is an ordinary stream operation
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&h_out, n_eval * sing::kDiffOut * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_diff();
        return fail(SING_ERR_NOMEM, "hipMalloc of differentiable-render buffers failed: %s", hipGetErrorString(e));
    }
    diff.P = std::uint32_t(P);
    diff.Pimg = std::uint32_t(Pimg);
    diff.n_eval = std::uint32_t(n_eval);
    diff.ncp = std::uint32_t(ncp);
    diff.C = std::uint32_t(C);
    diff_P = Pimg;
    diff_n = n_eval;
    diff_ncp = ncp;
    diff_C = C;
    return SING_OK;
}

namespace {

// The DiffSet of one parameter point (host): the scene in double, its
// (a/M, inclination) tangents as Dual<2>, and both rounded to float.
// r_in: given, or the prograde ISCO of the spin (drin_da its derivative).
sing::DiffSet make_diff_set(const SingUniforms& base, const SingDiffParams& p, const SingDiffCamera& cam, double w,
                            double r_in) {
    sing::DiffSet ds{};
    using D2 = sing::Dual<2, double>;
    const sing::DiffScene<D2> td = sing::make_diff_scene<D2>(base, D2::var(p.spin, 0), D2::var(p.incl, 1), D2(r_in),
                                                             D2(p.r_out), D2(p.t_scale), cam.distance, cam.azimuth,
                                                             cam.fov_y_deg, w);
    ds.td = td;
    ds.sd = sing::make_diff_scene<double>(base, p.spin, p.incl, r_in, p.r_out, p.t_scale, cam.distance, cam.azimuth,
                                          cam.fov_y_deg, w);
    // float copies are the double values rounded (same discrete model)
    auto to_f = [](const sing::DiffScene<double>& d) {
        sing::DiffScene<float> f{};
        for (int i = 0; i < 3; ++i) {
            f.fwd[i] = float(d.fwd[i]); f.right[i] = float(d.right[i]); f.up[i] = float(d.up[i]);
            f.rhat[i] = float(d.rhat[i]); f.that[i] = float(d.that[i]); f.phat[i] = float(d.phat[i]);
        }
        f.tan_half_fov = float(d.tan_half_fov); f.aspect = float(d.aspect); f.width_f = float(d.width_f);
        f.height_f = float(d.height_f); f.cam_r = float(d.cam_r); f.cam_theta = float(d.cam_theta);
        f.cam_phi = float(d.cam_phi); f.uphi_den = float(d.uphi_den); f.g_tt = float(d.g_tt); f.g_tp = float(d.g_tp);
        f.g_rr = float(d.g_rr); f.g_thth = float(d.g_thth); f.g_pp = float(d.g_pp); f.M = float(d.M); f.a = float(d.a);
        f.r_in = float(d.r_in); f.r_out = float(d.r_out); f.peak_T = float(d.peak_T); f.f_obs = float(d.f_obs);
        f.turbulence = float(d.turbulence); f.time_sec = float(d.time_sec); f.w = float(d.w); f.h = float(d.h);
        f.escape_r = float(d.escape_r); f.cut = float(d.cut);
        f.width = d.width; f.height = d.height; f.flags = d.flags; f.max_steps = d.max_steps;
        return f;
    };
    ds.sf = to_f(ds.sd);
    {
        // Dual<2, float> scene: values and tangents rounded from the double duals
        const D2* src = &td.fwd[0];
        sing::Dual<2, float>* dst = &ds.tf.fwd[0];
        constexpr std::size_t n_fields = (offsetof(sing::DiffScene<D2>, width)) / sizeof(D2);
        for (std::size_t i = 0; i < n_fields; ++i) {
            dst[i].v = float(src[i].v);
            dst[i].d[0] = float(src[i].d[0]);
            dst[i].d[1] = float(src[i].d[1]);
        }
        ds.tf.width = td.width; ds.tf.height = td.height; ds.tf.flags = td.flags; ds.tf.max_steps = td.max_steps;
    }
    ds.tscale = p.t_scale;
    ds.T0 = base.disc_peak_T;
    return ds;
}

// n parameter points (cameras: one per point, or n_cams == 1 for all) of one
// image size, evaluated in one set of launches.
int diff_run(SingContext* c, const SingUniforms* base, std::uint32_t n, const SingDiffParams* params,
             const SingDiffCamera* cams, std::uint32_t n_cams, const SingDiffOpts* opts_in, const float* target,
             const float* weights, bool backward, double* loss, double* grad, float* hdr_out, SingDiffStats* stats,
             const char* who, const float* cot = nullptr) {
    if (!c) return SING_ERR_ARG;
    c->complete_pending();
    if (!base || !params || !cams) return c->fail(SING_ERR_ARG, "%s: null argument", who);
    if (n < 1 || n > SING_DIFF_MAX_BATCH)
        return c->fail(SING_ERR_ARG, "%s: evaluations must be in [1, %u]", who, unsigned(SING_DIFF_MAX_BATCH));
    if (n_cams != 1 && n_cams != n) return c->fail(SING_ERR_ARG, "%s: give one camera or one per evaluation", who);
    if (base->width != c->width || base->height != c->height)
        return c->fail(SING_ERR_ARG, "%s: base size %ux%u != context %ux%u", who, base->width, base->height,
                       c->width, c->height);
    SingDiffOpts o{};
    if (opts_in) o = *opts_in;
    if (o.precision > SING_PREC_MIXED) return c->fail(SING_ERR_ARG, "%s: unknown precision %u", who, o.precision);
    const double w = o.soft_edge > 0.0 ? o.soft_edge : 1.0;
    const std::size_t C = o.checkpoint_every ? o.checkpoint_every : 64u;
    if (C > 4096) return c->fail(SING_ERR_ARG, "%s: checkpoint_every too large", who);
    if (o.r_in_isco > 1u) return c->fail(SING_ERR_ARG, "%s: r_in_isco must be 0 or 1", who);
    if (backward && !target && !cot) return c->fail(SING_ERR_ARG, "%s: null %s", who, "target");
    if (o.target_per_eval > 1u) return c->fail(SING_ERR_ARG, "%s: target_per_eval must be 0 or 1", who);
    if (o.target_per_eval && (cot || !target))
        return c->fail(SING_ERR_ARG, "%s: target_per_eval needs the target images", who);
    const std::size_t Pimg = std::size_t(c->width) * c->height;
    if (Pimg * n >= (std::size_t(1) << 31)) return c->fail(SING_ERR_ARG, "%s: too many pixels in one call", who);
    std::vector<sing::DiffSet> sets(n);
    std::vector<double> drin_da(n, 0.0);
    for (std::uint32_t k = 0; k < n; ++k) {
        const SingDiffParams& p = params[k];
        const SingDiffCamera& cam = cams[n_cams == 1 ? 0 : k];
        if (!(p.spin >= 0.0 && p.spin < 1.0)) return c->fail(SING_ERR_ARG, "%s: spin must be in [0, 1)", who);
        if (!(p.t_scale > 0.0) || !(cam.distance > 0.0))
            return c->fail(SING_ERR_ARG, "%s: need t_scale > 0, distance > 0", who);
        double r_in = p.r_in;
        if (o.r_in_isco) {
            const sing::Dual<1, double> ri = sing::isco_prograde(sing::Dual<1, double>::var(p.spin, 0));
            r_in = ri.v;
            drin_da[k] = ri.d[0];
        }
        if (!(p.r_out > r_in)) return c->fail(SING_ERR_ARG, "%s: need r_out > r_in (r_in = %g)", who, r_in);
        sets[k] = make_diff_set(*base, p, cam, w, r_in);
    }
    const std::size_t ncp = base->max_steps / C + 1;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hip_fail(e, "hipSetDevice");
    int rc = c->ensure_diff(Pimg, n, ncp, C);
    if (rc != SING_OK) return rc;
    if ((rc = c->ensure_events(6)) != SING_OK) return rc;
    if ((rc = c->ensure_join_events(4)) != SING_OK) return rc;
    const std::size_t P = Pimg * n;

    if ((e = hipMemcpyAsync(c->d_dset, sets.data(), n * sizeof(sing::DiffSet), hipMemcpyHostToDevice, c->stream)) !=
        hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(scene)");
    const std::size_t n_img = o.target_per_eval ? n : 1u;  // target (and weight) images
    if (target || cot) {  // the cotangent image travels in the target buffer
        if ((e = hipMemcpyAsync(c->d_target, cot ? cot : target, n_img * Pimg * 3 * sizeof(float),
                                hipMemcpyHostToDevice, c->stream)) != hipSuccess)
            return c->hip_fail(e, "hipMemcpyAsync(target)");
    } else if ((e = hipMemsetAsync(c->d_target, 0, Pimg * 3 * sizeof(float), c->stream)) != hipSuccess) {
        return c->hip_fail(e, "hipMemsetAsync(target)");
    }
    if (weights && !cot &&
        (e = hipMemcpyAsync(c->d_weight, weights, n_img * Pimg * sizeof(float), hipMemcpyHostToDevice,
                            c->stream)) !=
            hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(weights)");
    if ((e = hipMemsetAsync(c->diff.count2, 0, n * sizeof(std::uint32_t), c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemsetAsync(count2)");
    if ((e = hipMemsetAsync(c->diff.count, 0, n * sizeof(std::uint32_t), c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemsetAsync(count)");
    if (!backward && (e = hipMemsetAsync(c->diff.grad, 0, P * 5 * sizeof(double), c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemsetAsync(grad)");
    sing::DiffBuffers b = c->diff;
    b.target = c->d_target;
    b.weight = (weights && !cot) ? c->d_weight : nullptr;
    b.cot = cot ? c->d_target : nullptr;
    b.tstride = o.target_per_eval ? std::uint32_t(Pimg) : 0u;
    (void)hipEventRecord(c->events[0], c->stream);
    if ((e = sing::launch_diff(c->d_dset, b, o.precision, backward, c->d_out, c->stream, c->stream2,
                               c->events[1], c->join_events.data())) !=
        hipSuccess)
        return c->hip_fail(e, "differentiable render launch");
    (void)hipEventRecord(c->events[5], c->stream);
    // losses, gradients, step totals and the FP64 list length in one copy to
    // pinned memory (the step totals are reduced on the device)
    if ((e = hipMemcpyAsync(c->h_out, c->d_out, n * sing::kDiffOut * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(result)");
    if (hdr_out &&
        (e = hipMemcpyAsync(hdr_out, c->diff.hdr, P * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream)) !=
            hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(hdr)");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
    std::uint64_t steps = 0, listed64 = 0;
    for (std::uint32_t k = 0; k < n; ++k) {
        const double* out = c->h_out + std::size_t(k) * sing::kDiffOut;
        listed64 += std::uint64_t(out[sing::kReduceVals]);
        if (loss) loss[k] = out[0];
        if (grad) {
            double* g = grad + 5u * k;
            for (int j = 0; j < 5; ++j) g[j] = out[1 + j];
            g[0] += g[2] * drin_da[k];
        }
        steps += std::uint64_t(out[6]);
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float a = 0.0f, bms = 0.0f;
        (void)hipEventElapsedTime(&a, c->events[0], c->events[1]);
        (void)hipEventElapsedTime(&bms, c->events[1], c->events[5]);
        stats->forward_ms = a;
        stats->backward_ms = bms;
        stats->rays = P;
        stats->steps = steps;
        stats->fp64_rays = (o.precision == SING_PREC_FP64) ? P : listed64;
    }
    return SING_OK;
}

}  // namespace

extern "C" {

int sing_diff_forward(SingContext* ctx, const SingUniforms* base, const SingDiffParams* p, const SingDiffCamera* cam,
                      const SingDiffOpts* opts, float* hdr_out, SingDiffStats* stats) {
    return diff_run(ctx, base, 1, p, cam, 1, opts, nullptr, nullptr, false, nullptr, nullptr, hdr_out, stats,
                    "sing_diff_forward");
}

int sing_diff_loss_grad(SingContext* ctx, const SingUniforms* base, const SingDiffParams* p,
                        const SingDiffCamera* cam, const SingDiffOpts* opts, const float* target,
                        const float* weights, double* loss, double* grad, float* hdr_out, SingDiffStats* stats) {
    return diff_run(ctx, base, 1, p, cam, 1, opts, target, weights, true, loss, grad, hdr_out, stats,
                    "sing_diff_loss_grad");
}

int sing_diff_loss_grad_batch(SingContext* ctx, const SingUniforms* base, uint32_t n, const SingDiffParams* params,
                              const SingDiffCamera* cams, uint32_t n_cams, const SingDiffOpts* opts,
                              const float* target, const float* weights, double* loss, double* grad, float* hdr_out,
                              SingDiffStats* stats) {
    return diff_run(ctx, base, n, params, cams, n_cams, opts, target, weights, true, loss, grad, hdr_out, stats,
                    "sing_diff_loss_grad_batch");
}

int sing_diff_vjp(SingContext* ctx, const SingUniforms* base, const SingDiffParams* p, const SingDiffCamera* cam,
                  const SingDiffOpts* opts, const float* dl_dhdr, double* pairing, double* grad, float* hdr_out,
                  SingDiffStats* stats) {
    if (ctx && !dl_dhdr) return ctx->fail(SING_ERR_ARG, "sing_diff_vjp: null dl_dhdr");
    return diff_run(ctx, base, 1, p, cam, 1, opts, nullptr, nullptr, true, pairing, grad, hdr_out, stats,
                    "sing_diff_vjp", dl_dhdr);
}

int sing_diff_capture_pixels(SingContext* c, double* grad_out, uint32_t* meta_out) {
    if (!c) return SING_ERR_ARG;
    c->complete_pending();
    if (!c->diff_P || !c->d_dset) return c->fail(SING_ERR_STATE, "sing_diff_capture_pixels: no differentiable frame");
    hipError_t e = hipSetDevice(c->device);
    const std::size_t P = c->diff_P;  // one image: the first evaluation of the last call
    if (e == hipSuccess && grad_out) e = hipMemcpy(grad_out, c->diff.grad, P * 5 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && meta_out) e = hipMemcpy(meta_out, c->diff.meta, P * 4 * sizeof(std::uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "sing_diff_capture_pixels");
    return SING_OK;
}

// ---- diagnostics: the adjoint math evaluated on the host
void sing_diag_kerr_rhs(const double* y, const double* c, double* out) {
    const sing::KY<double> d = sing::kerr_rhs<double>({y[0], y[1], y[2], y[3], y[4]}, {c[0], c[1], c[2], c[3]});
    out[0] = d.r; out[1] = d.th; out[2] = d.ph; out[3] = d.pr; out[4] = d.pt;
}
void sing_diag_kerr_rhs_vjp(const double* y, const double* c, const double* v, double* gy, double* gc) {
    sing::KY<double> g = {0, 0, 0, 0, 0};
    sing::KG<double> k = {0, 0, 0};
    sing::kerr_rhs_vjp<double>({y[0], y[1], y[2], y[3], y[4]}, {c[0], c[1], c[2], c[3]}, {v[0], v[1], v[2], v[3], v[4]},
                               g, k);
    gy[0] = g.r; gy[1] = g.th; gy[2] = g.ph; gy[3] = g.pr; gy[4] = g.pt;
    gc[0] = k.E; gc[1] = k.L; gc[2] = k.a;
}
void sing_diag_kerr_step(const double* y, const double* c, double h, double* out) {
    const sing::KY<double> d = sing::kerr_step<double>({y[0], y[1], y[2], y[3], y[4]}, {c[0], c[1], c[2], c[3]}, h);
    out[0] = d.r; out[1] = d.th; out[2] = d.ph; out[3] = d.pr; out[4] = d.pt;
}
void sing_diag_kerr_step_vjp(const double* y, const double* c, double h, const double* lam, double* gy, double* gc) {
    sing::KG<double> k = {0, 0, 0};
    const sing::KY<double> g = sing::kerr_step_vjp<double>({y[0], y[1], y[2], y[3], y[4]}, {c[0], c[1], c[2], c[3]},
                                                          h, {lam[0], lam[1], lam[2], lam[3], lam[4]}, k);
    gy[0] = g.r; gy[1] = g.th; gy[2] = g.ph; gy[3] = g.pr; gy[4] = g.pt;
    gc[0] = k.E; gc[1] = k.L; gc[2] = k.a;
}

}  // extern "C"
