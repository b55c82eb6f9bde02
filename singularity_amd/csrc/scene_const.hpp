// singularity_amd/csrc/scene_const.hpp — per-scene constants of the trace
// kernels.  Everything here depends only on the 160-byte uniform block and the
// render options, so the host evaluates it once per scene (in float for the
// FP32 path, in double for the FP64 path) instead of once per sample as the
// reference kernel does (render/cuda/kernels/geodesic_kernel.cu:188-223,
// 303-349).  Shared by host (.cpp/.hip host code) and device code.
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/sing_abi.h"

namespace sing {

constexpr std::uint32_t kTermLive = 0;      // ray still being integrated
constexpr std::uint32_t kTermHorizon = 1;   // -> SING_CLASS_HORIZON
constexpr std::uint32_t kTermEscape = 2;    // -> SING_CLASS_ESCAPE
constexpr std::uint32_t kTermDisc = 3;      // -> SING_CLASS_DISC
constexpr std::uint32_t kTermBudget = 4;    // -> SING_CLASS_BUDGET
constexpr std::uint32_t kTermInitFail = 5;  // -> SING_CLASS_INITFAIL

template <class R> struct SceneConst {
    // camera basis and projection (geodesic_kernel.cu:188-199)
    R fwd[3], right[3], up[3];
    R tan_half_fov, aspect, width_f, height_f;
    // camera position in Boyer-Lindquist coordinates and its local frame (:202-216)
    R cam_r, cam_theta, cam_phi, sin_th, cos_th;
    R rhat[3], that[3], phat[3];
    R uphi_den;  // cam_r * max(sin_th, 1e-6)
    // metric, integrator and termination
    R M, a, a2, twoM, rs;
    R h, escape_r, cut;  // cut: Schwarzschild horizon_cut, Kerr 1.02 r+ (geodesic_kernel.cu:349)
    // Kerr metric at the camera (geodesic_kernel.cu:306-319)
    R g_tt, g_tp, g_rr, g_thth, g_pp, q_cos2, q_inv_s2;
    R f_obs;  // Kerr shading: 1 - 2M/cam_r
    R f_cam;  // Schwarzschild: 1 - rs/cam_r
    // disc and shading
    R r_in, r_out, peak_T, turbulence, time_sec, exposure, ph1, ph2;
    // mixed precision: rays whose photon-orbit margin |R(r*)|/r*^4 < band go FP64
    R band;
    // error-controlled stepping (Kerr): per-step tolerance, 0 = the reference's fixed step
    R tol;
    std::uint32_t width, height, metric, flags, max_steps, spp;
    // SING_REFINE_* (fixed-step Kerr): FP64 crossing root, null-constraint re-projection
    std::uint32_t refine;
    // grid sampler: spp = ss * ss offsets from the grid table, ACES of the mean radiance
    std::uint32_t grid;
};

// Same arithmetic as the reference kernel's per-sample camera and metric
// evaluation, done once.  R = float reproduces the FP32 values; R = double
// evaluates the same formulas in double from the float uniforms.
template <class R> inline SceneConst<R> make_scene_const(const SingUniforms& u, const SingRenderOpts& o) {
    using std::acos;
    using std::atan2;
    using std::cos;
    using std::sin;
    using std::sqrt;
    SceneConst<R> s{};
    s.fwd[0] = u.cam_fwd.x; s.fwd[1] = u.cam_fwd.y; s.fwd[2] = u.cam_fwd.z;
    s.right[0] = u.cam_right.x; s.right[1] = u.cam_right.y; s.right[2] = u.cam_right.z;
    s.up[0] = u.cam_up.x; s.up[1] = u.cam_up.y; s.up[2] = u.cam_up.z;
    s.tan_half_fov = u.tan_half_fov;
    s.aspect = u.aspect;
    s.width_f = R(u.width);
    s.height_f = R(u.height);
    const R px = u.cam_pos.x, py = u.cam_pos.y, pz = u.cam_pos.z;
    s.cam_r = sqrt(px * px + py * py + pz * pz);
    R zc = pz / std::fmax(s.cam_r, R(1e-6));
    zc = std::fmin(std::fmax(zc, R(-1)), R(1));
    s.cam_theta = acos(zc);
    s.cam_phi = atan2(py, px);
    s.sin_th = sin(s.cam_theta);
    s.cos_th = cos(s.cam_theta);
    const R cp = cos(s.cam_phi), sp = sin(s.cam_phi);
    s.rhat[0] = s.sin_th * cp; s.rhat[1] = s.sin_th * sp; s.rhat[2] = s.cos_th;
    s.that[0] = s.cos_th * cp; s.that[1] = s.cos_th * sp; s.that[2] = -s.sin_th;
    s.phat[0] = -sp; s.phat[1] = cp; s.phat[2] = R(0);
    s.uphi_den = s.cam_r * std::fmax(s.sin_th, R(1e-6));

    s.M = u.mass_M;
    s.a = u.spin_a;
    s.a2 = s.a * s.a;
    s.twoM = R(2) * s.M;
    s.rs = u.rs;
    s.h = u.h_step;
    s.escape_r = u.escape_r;
    const R M = s.M, a = s.a, cr = s.cam_r;
    if (u.metric_type == SING_METRIC_KERR) {
        s.cut = R(1.02) * (M + sqrt(std::fmax(M * M - a * a, R(0))));
    } else {
        s.cut = u.horizon_cut;
    }
    const R cth = s.cos_th;  // geodesic_kernel.cu:307 recomputes cosf(cam_theta): same value
    const R Sigma_c = cr * cr + a * a * cth * cth;
    const R Delta_c = cr * cr - R(2) * M * cr + a * a;
    const R s2c = s.sin_th * s.sin_th;
    s.g_tt = -(R(1) - R(2) * M * cr / Sigma_c);
    s.g_tp = R(-2) * M * cr * a * s2c / Sigma_c;
    s.g_rr = Sigma_c / Delta_c;
    s.g_thth = Sigma_c;
    s.g_pp = s2c * ((cr * cr + a * a) * (cr * cr + a * a) - a * a * Delta_c * s2c) / Sigma_c;
    s.q_cos2 = cth * cth;
    s.q_inv_s2 = R(1) / std::fmax(s2c, R(1e-6));
    s.f_obs = R(1) - R(2) * M / cr;
    s.f_cam = R(1) - s.rs / cr;

    s.r_in = u.disc_r_inner;
    s.r_out = u.disc_r_outer;
    const R scale = o.peak_T_scale > 0.0f ? R(o.peak_T_scale) : R(1);
    s.peak_T = R(u.disc_peak_T) * scale;
    s.turbulence = u.disc_turbulence;
    s.time_sec = u.time_sec;
    s.exposure = u.exposure;
    s.ph1 = R(0);
    s.ph2 = R(0);
    if (o.turb_seed != 0) {
        // SplitMix64 (state = seed): the first two outputs' top 53 bits -> two phases in [0, 2 pi)
        std::uint64_t z = std::uint64_t(o.turb_seed);
        auto next = [&z]() {
            std::uint64_t x = (z += 0x9E3779B97F4A7C15ull);
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            return x ^ (x >> 31);
        };
        const double twopi = 6.283185307179586476925286766559;
        s.ph1 = R(double(next() >> 11) * (1.0 / 9007199254740992.0) * twopi);
        s.ph2 = R(double(next() >> 11) * (1.0 / 9007199254740992.0) * twopi);
    }
    s.band = o.band_width > 0.0f ? R(o.band_width) : R(0.03);
    s.tol = o.adapt_tol > 0.0f ? R(o.adapt_tol) : R(0);
    s.width = u.width;
    s.height = u.height;
    s.metric = u.metric_type;
    s.flags = u.flags;
    s.max_steps = u.max_steps;
    std::uint32_t spp = u.supersample;
    s.spp = spp < 1u ? 1u : (spp > 1024u ? 1024u : spp);
    s.refine = o.refine;
    s.grid = o.sampler == SING_SAMPLER_GRID ? 1u : 0u;
    if (s.grid) {
        const std::uint32_t ss = s.spp < 4u ? s.spp : 4u;
        s.spp = ss * ss;
    }
    return s;
}

// Grid sampler offsets (render/metal/shaders/geodesic_kernel.metal:417-424):
// sample sy * ss + sx sits at ((sx + 0.5) / ss, (sy + 0.5) / ss), in float.
// The tables for ss = 1..4 are stored back to back; grid_table_base gives where
// the table for one ss starts.
constexpr std::uint32_t kGridTableLen = 1u + 4u + 9u + 16u;
inline std::uint32_t grid_table_base(std::uint32_t ss) { return (ss - 1u) * ss * (2u * ss - 1u) / 6u; }
inline void grid_offset_f(std::uint32_t s, std::uint32_t ss, float& ox, float& oy) {
    const float sub_w = 1.0f / float(ss);
    ox = (float(s % ss) + 0.5f) * sub_w;
    oy = (float(s / ss) + 0.5f) * sub_w;
}

// Halton radical inverse in float (geodesic_kernel.cu:442-451); the kernels
// read a host-built table of (halton(s+1,2), halton(s+1,3)).
inline float halton_f(std::uint32_t i, std::uint32_t base) {
    float f = 1.0f, r = 0.0f;
    while (i > 0u) {
        f /= float(base);
        r += f * float(i % base);
        i /= base;
    }
    return r;
}

}  // namespace sing
