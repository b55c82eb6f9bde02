// singularity_amd/csrc/trace_core.hpp — per-ray device code shared by the
// frame pipeline (trace.hip) and the batched pipeline (batch.hip): the ray
// state, termination checks + one (error-controlled) RK4 step, camera
// initialisation, the termination record, the work-item mapping and the
// output class of a record.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "physics.hpp"
#include "trace.hpp"

namespace sing {
namespace core {

template <class S> struct RayState {
    S r, th, ph;
    S v0, v1, v2, v3;  // Kerr: p_r, p_theta, E, L_z.  Schwarzschild: u^t, u^r, u^theta, u^phi
    S prev_r, prev_th, prev_uph;  // prev_uph: Schwarzschild u^phi; error-controlled Kerr: phi
    S prev_v0, prev_v1, hprev;    // error-controlled Kerr: p_r, p_theta before the step, its size
    S aux0, aux1;      // disc crossing radius; u^phi (Schwarzschild) or phi (error-controlled Kerr) at it
    S near_eq;         // mixed precision: min |theta - pi/2| over the steps (0 = ambiguous crossing)
    S hcur;            // error-controlled stepping: current step size
    S sth, cth;        // fixed-step Kerr: sin and cos of th carried from the previous step
    std::uint32_t steps;
    std::uint32_t tries;  // error-controlled stepping: rejections of the current step
};

// Mixed-precision deferral thresholds.  Comparing FP32 and FP64 traces of the
// same rays (57 600 rays of 64 varied Kerr scenes and the a = 0.94 4K scene),
// outcomes or radiance differ (> 1e-3) for rays near a discontinuity of the
// discrete algorithm — a state within kGrazeTheta of the equator (the crossing
// test flips), a crossing within kEdgeRel of an annulus edge, an escape within
// kStarCell cells of a starfield boundary next to a star (tested during and
// after the FP32 march) — and for rays whose trajectory error is amplified:
// close to a photon orbit, reaching 2 r+, or reaching 0.03 rad of the polar
// axis (decided before the march from the ray's constants, hard_ray in
// physics.hpp).  Such rays are traced in FP64.
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

// Per-ray products of E and L_z for the Kerr RHS (E, L_z are constants of
// the motion, so these are formed once per batch of steps, not per step).
template <class R, class S, bool KERR>
__device__ __forceinline__ KerrK<R> ray_k(const RayState<S>& y, const SceneConst<R>& sc) {
    if constexpr (KERR) return kerr_k(sc, R(y.v2), R(y.v3));
    else return KerrK<R>{};
}

// Equator crossing of a fixed step with the root refined in FP64
// (SING_REFINE_CROSS64): the cubic Hermite dense output of the step from its
// end states and their RHS derivatives (in the march's precision R; the step
// is the damped one the march took, kept in hprev), then, widened to double,
// three Newton iterations on theta(s) = pi/2 from the linear fraction and r,
// phi at the root.  The derivatives' rounding moves the root by ~1 ulp of s;
// the FP32 subtraction theta(s) - pi/2 would move it by ~1e-7 / |dtheta|.
template <class R, class S>
__device__ __forceinline__ void refined_crossing(const RayState<S>& y, const SceneConst<R>& sc, const KerrK<R>& k,
                                                 R& rc, R& phc) {
    using W = double;
    const KerrDense<R> d = kerr_dense<R>(R(y.prev_r), R(y.prev_th), R(y.prev_uph), R(y.prev_v0), R(y.prev_v1),
                                         R(y.r), R(y.th), R(y.ph), R(y.v0), R(y.v1), R(y.hprev), sc, k);
    const W mid = W(1.5707963267948966192313216916397514);
    const W th0 = W(d.th0), th1 = W(d.th1);
    const W s = dense_root<W>(th0, W(d.dth0), th1, W(d.dth1), mid, (mid - th0) / (th1 - th0));
    rc = R(hermite<W>(s, W(d.r0), W(d.dr0), W(d.r1), W(d.dr1)));
    phc = R(hermite<W>(s, W(d.ph0), W(d.dph0), W(d.ph1), W(d.dph1)));
}

// Termination checks in the reference's order (geodesic_kernel.cu:351-401)
// followed by one RK4 step.  Arithmetic in R, storage in S; k = ray_k(y, sc)
// (Kerr).  With ADAPT (Kerr) the step is one attempt of the error-controlled
// step; a rejected attempt leaves the state unchanged (the checks it repeats
// are idempotent) and counts in *rejected.  The four exit conditions are
// tested together so that a marching ray takes one branch; which one fired
// (budget, horizon, escape, then a crossing inside the annulus) is sorted out
// only on the rare exit path.  XTRA (fixed-step Kerr): the step keeps the
// state before it for the crossing refinement, and sc.refine selects the FP64
// crossing root and the null-constraint re-projection.  phase = y.steps mod 8
// (fixed-step Kerr): run_steps passes its loop counter, the same for every
// live lane of the wave (see there), so the branch on it is uniform.
template <class R, class S, bool KERR, bool TRACK = false, bool ADAPT = false, bool XTRA = false>
__device__ __forceinline__ std::uint32_t advance(RayState<S>& y, const SceneConst<R>& sc, const KerrK<R>& k,
                                                 std::uint32_t* rejected, std::uint32_t phase) {
    R r = R(y.r), th = R(y.th);
    const R mid = R(1.5707963267948966192313216916397514);
    const R dp = R(y.prev_th) - mid, dc = th - mid;
    const bool budget = y.steps >= sc.max_steps;
    const bool horizon = !__builtin_isfinite(r) || r < sc.cut;
    const bool escape = r > sc.escape_r;
    const bool cross = (sc.flags & SING_FLAG_DISC_ON) && dp * dc < R(0);
    if (__builtin_expect(budget || horizon || escape || cross, false)) {
        if (budget) return kTermBudget;
        if (horizon) return kTermHorizon;
        if (escape) {
            if (ADAPT) {
                R the = th, phe = R(y.ph);
                if (y.steps > 0u) {
                    const KerrDense<R> d = kerr_dense(R(y.prev_r), R(y.prev_th), R(y.prev_uph), R(y.prev_v0),
                                                      R(y.prev_v1), r, th, R(y.ph), R(y.v0), R(y.v1), R(y.hprev), sc,
                                                      k);
                    kerr_dense_escape(d, sc.escape_r, the, phe);
                }
                y.aux0 = S(the);
                y.aux1 = S(phe);
            }
            return kTermEscape;
        }
        // a crossing of the equatorial plane (disc_intersection.h:23-55)
        const R pth = R(y.prev_th);
        const R frac = (mid - pth) / (th - pth);
        const R pr0 = R(y.prev_r);
        R rc = pr0 + frac * (r - pr0);
        R phc = R(0);
        if (ADAPT) {
            const KerrDense<R> d = kerr_dense(pr0, pth, R(y.prev_uph), R(y.prev_v0), R(y.prev_v1), r, th, R(y.ph),
                                              R(y.v0), R(y.v1), R(y.hprev), sc, k);
            kerr_dense_crossing(d, frac, rc, phc);
        }
        if (XTRA) {
            phc = R(y.ph);  // the reference shades at phi of the step after the crossing
            if ((sc.refine & SING_REFINE_CROSS64) && y.steps > 0u) refined_crossing(y, sc, k, rc, phc);
        }
        if (TRACK) {
            const R tol = R(kEdgeRel) * rc;
            const R d_in = rc - sc.r_in, d_out = rc - sc.r_out;
            if ((d_in < tol && d_in > -tol) || (d_out < tol && d_out > -tol)) y.near_eq = S(0);
        }
        if (rc >= sc.r_in && rc <= sc.r_out) {
            y.aux0 = S(rc);
            if (ADAPT || XTRA) y.aux1 = S(phc);
            if (!KERR) {
                const R pu = R(y.prev_uph);
                y.aux1 = S(pu + frac * (R(y.v3) - pu));
            }
            return kTermDisc;
        }
    }
    y.prev_th = y.th;
    y.prev_r = y.r;
    R ph = R(y.ph);
    if (KERR) {
        R pr = R(y.v0), pt = R(y.v1);
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
            // sin/cos of theta: evaluated in full on every 8th step of the ray
            // (steps = 0 mod 8, so the result does not depend on the launch
            // shape), otherwise carried from the previous step, rotated by the
            // exact difference of the stored angles (sincos_shift); the
            // rotation error stays below a few ulp between full evaluations.
            R st, ct;
            if (phase == 0u) {
                dsincos(th, &st, &ct);
            } else {
                st = R(y.sth);
                ct = R(y.cth);
            }
            const R th0 = th;
            const R hs = sc.h * kerr_pole_scale(st);
            if (XTRA) {
                y.prev_uph = y.ph;
                y.prev_v0 = y.v0;
                y.prev_v1 = y.v1;
                y.hprev = S(hs);
            }
            kerr_rk4_raw<R>(r, th, ph, pr, pt, hs, st, ct, sc, k);
            if (XTRA) {
                sincos_shift(th0, th - th0, st, ct, &st, &ct);
                if (sc.refine & SING_REFINE_REPROJECT) kerr_reproject(r, pr, pt, st, sc, k);
                y.sth = S(st);
                y.cth = S(ct);
            } else if (phase != 7u) {
                sincos_shift(th0, th - th0, st, ct, &st, &ct);
                y.sth = S(st);
                y.cth = S(ct);
            }
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
    if (TRACK) y.near_eq = S(pos_min(R(y.near_eq), dabs(th - mid)));
    return kTermLive;
}

// Up to n steps of a live ray (advance); returns its termination code, or
// kTermLive if it is still marching, and clears live when it ends.  n is a
// multiple of 8 and every ray starts (steps = 0) at a call's first step, so at
// step i of the call each live lane has taken i mod 8 steps modulo 8 (a
// fixed-step ray advances one step per iteration until it ends).
template <class R, class S, bool KERR, bool TRACK, bool ADAPT, bool XTRA = false>
__device__ __forceinline__ std::uint32_t run_steps(RayState<S>& y, const SceneConst<R>& sc, bool& live,
                                                   std::uint32_t n, std::uint32_t* rejected) {
    std::uint32_t term = kTermLive;
    const KerrK<R> k = ray_k<R, S, KERR>(y, sc);
    for (std::uint32_t i = 0; i < n; ++i) {
        if (live) {
            const std::uint32_t t = advance<R, S, KERR, TRACK, ADAPT, XTRA>(y, sc, k, rejected, i & 7u);
            if (t != kTermLive) {
                live = false;
                term = t;
            }
        }
    }
    return term;
}

// Mixed precision, after the FP32 march: does the outcome need FP64?  (A
// grazing crossing or an annulus-edge crossing; the trajectory-amplification
// cases are caught before the march by hard_ray.)
template <class S>
__device__ __forceinline__ bool needs_fp64(const RayState<S>& y, std::uint32_t) {
    return float(y.near_eq) < kGrazeTheta;
}

// Ray initialisation from the camera (geodesic_kernel.cu:229-242, 303-347).
// Returns false on a Kerr initialisation failure (no real root at the camera).
// With hard_out, also decides from the ray's constants whether FP32 is not
// accurate enough for it (hard_ray).
template <class R, class S, bool KERR>
__device__ __forceinline__ bool init_ray(RayState<S>& y, const SceneConst<R>& sc, std::uint32_t gx, std::uint32_t gy,
                                         R ox, R oy, bool* hard_out, std::uint32_t* kind_out = nullptr) {
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
        if (hard_out) {
            const R Q = pt * pt + sc.q_cos2 * (L * L * sc.q_inv_s2 - sc.a2 * E * E);
            const R rp = sc.M + dsqrt(dmax(sc.M * sc.M - sc.a2, R(0)));
            const R iE = R(1) / E;
            if (kind_out) {
                *kind_out = hard_ray_kind<R>(L * iE, Q * iE * iE, sc.a, sc.M, rp, ur < R(0), sc.band);
                *hard_out = *kind_out != 0u;
            } else {
                *hard_out = hard_ray<R>(L * iE, Q * iE * iE, sc.a, sc.M, rp, ur < R(0), sc.band);
            }
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
        if (hard_out) {
            // Schwarzschild: xi = L_z/E, eta = (L^2 - L_z^2)/E^2 (Carter's constant at a = 0)
            const R E = sc.f_cam * ut;
            const R L2 = cr * cr * cr * cr * (uth * uth + sc.sin_th * sc.sin_th * uph * uph);
            const R Lz = cr * cr * sc.sin_th * sc.sin_th * uph;
            const R iE = R(1) / E;
            if (kind_out) {
                *kind_out = hard_ray_kind<R>(Lz * iE, (L2 - Lz * Lz) * iE * iE, R(0), sc.M, sc.twoM, ur < R(0), sc.band);
                *hard_out = *kind_out != 0u;
            } else {
                *hard_out = hard_ray<R>(Lz * iE, (L2 - Lz * Lz) * iE * iE, R(0), sc.M, sc.twoM, ur < R(0), sc.band);
            }
        }
    }
    return true;
}

// A ray set up by the set-up kernel: the camera position plus the initial
// (v0, v1, v2, v3) that init_ray computed for it.
template <class S, bool KERR>
__device__ __forceinline__ void start_ray(RayState<S>& y, const SceneConst<float>& sc, const float4 v) {
    y.r = S(sc.cam_r);
    y.th = S(sc.cam_theta);
    y.ph = S(sc.cam_phi);
    y.prev_r = y.r;
    y.prev_th = y.th;
    y.steps = 0;
    y.near_eq = S(1e30);
    y.hcur = S(sc.h);
    y.tries = 0u;
    y.v0 = S(v.x);
    y.v1 = S(v.y);
    y.v2 = S(v.z);
    y.v3 = S(v.w);
    if (!KERR) y.prev_uph = y.v3;
}

// DISC_AUX: a disc record's phi is the one kept at the crossing (aux1).
template <class S, bool KERR, bool ADAPT = false, bool DISC_AUX = ADAPT>
__device__ __forceinline__ Rec make_rec(const RayState<S>& y, std::uint32_t term) {
    Rec rc;
    rc.meta = term | (y.steps << 4);
    if (term == kTermDisc) {
        rc.a = float(y.aux0);
        rc.b = DISC_AUX ? float(y.aux1) : float(y.ph);
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

// Placeholder record of a ray handed to the FP64 pass: its slot would otherwise
// hold a stale record until that pass overwrites it, and the audit kernel (which
// runs in between) must not mistake it for an escape.
__device__ __forceinline__ Rec deferred_marker() {
    Rec rc;
    rc.a = rc.b = rc.c = 0.0f;
    rc.meta = kRecDeferred;  // term bits = live
    return rc;
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

}  // namespace core
}  // namespace sing
