// singularity_amd/csrc/diff_scene.hpp — the differentiable Kerr scene.
//
// Parameters p = (a/M, inclination, r_in/M, r_out/M, peak-temperature scale).
// The camera is the reference's orbital camera (cuda_cli/main.cpp:68-109) at
// elevation pi/2 - inclination; the disc has soft annulus edges of width w
// (opacity ramps linearly over [r_in, r_in + w] and [r_out - w, r_out]), so the
// image is differentiable in r_in and r_out; the crossing's azimuth is
// interpolated like its radius, so the disc pattern does not jump when a
// crossing moves from one RK4 step to the next.  w -> 0 recovers the
// reference's hard annulus (shared_shader/disc_intersection.h:53-55), whose
// Kerr shading takes phi at the step after the crossing (geodesic_kernel.cu:380).
//
// Everything here is generic over the scalar: float / double for the march,
// Dual<2, T> for d/d(a, inclination) of the camera initialisation, Dual<6, T>
// for the emission Jacobian.  Host and device share the code.
#pragma once

#include <cstdint>

#include "../../include/sing_abi.h"
#include "adjoint_math.hpp"

namespace sing {

constexpr int kMaxCross = 8;  // partial (0 < alpha < 1) crossings kept per ray; the 8th is opaque

template <class R> struct DiffScene {
    R fwd[3], right[3], up[3];
    R tan_half_fov, aspect, width_f, height_f;
    R cam_r, cam_theta, cam_phi;
    R rhat[3], that[3], phat[3];
    R uphi_den;
    R g_tt, g_tp, g_rr, g_thth, g_pp;
    R M, a;
    R r_in, r_out, peak_T, f_obs, turbulence, time_sec, w;
    R h, escape_r, cut;
    std::uint32_t width, height, flags, max_steps;
};

// Scene from parameters.  base supplies resolution, integrator, flags, T0,
// turbulence, time and mass; p supplies (a/M, incl, r_in/M, r_out/M, Ts).
template <class R>
SING_HD DiffScene<R> make_diff_scene(const SingUniforms& base, R spin, R incl, R rin, R rout, R tscale, double dist,
                                     double az, double fov_deg, double soft_w) {
    DiffScene<R> s{};
    const R M = R(base.mass_M);
    const R elev = R(1.5707963267948966192313216916397514) - incl;
    const R d = R(dist);
    const R ce = hcos(elev), se = hsin(elev), ca = R(cos(az)), sa = R(sin(az));
    const R pos[3] = {d * ce * ca, d * ce * sa, d * se};
    const R inv = R(1) / d;
    const R f[3] = {-pos[0] * inv, -pos[1] * inv, -pos[2] * inv};
    R rt[3] = {f[1], -f[0], R(0)};
    const R rl = hsqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    const R ir = (rl > R(1e-6)) ? R(1) / rl : R(1);
    for (int i = 0; i < 3; ++i) rt[i] = rt[i] * ir;
    const R up[3] = {f[1] * rt[2] - f[2] * rt[1], f[2] * rt[0] - f[0] * rt[2], f[0] * rt[1] - f[1] * rt[0]};
    for (int i = 0; i < 3; ++i) {
        s.fwd[i] = f[i];
        s.right[i] = rt[i];
        s.up[i] = up[i];
    }
    s.tan_half_fov = htan(R(0.5) * R(fov_deg * (3.14159265358979323846 / 180.0)));
    s.width_f = R(double(base.width));
    s.height_f = R(double(base.height));
    s.aspect = R(base.aspect);
    // camera position in Boyer-Lindquist coordinates (geodesic_kernel.cu:202-216)
    s.cam_r = hsqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
    s.cam_theta = hacos(hclamp(pos[2] / hmax(s.cam_r, R(1e-6)), R(-1), R(1)));
    s.cam_phi = hatan2(pos[1], pos[0]);
    const R st = hsin(s.cam_theta), ct = hcos(s.cam_theta), cp = hcos(s.cam_phi), sp = hsin(s.cam_phi);
    s.rhat[0] = st * cp; s.rhat[1] = st * sp; s.rhat[2] = ct;
    s.that[0] = ct * cp; s.that[1] = ct * sp; s.that[2] = -st;
    s.phat[0] = -sp; s.phat[1] = cp; s.phat[2] = R(0);
    s.uphi_den = s.cam_r * hmax(st, R(1e-6));
    // Kerr metric at the camera (geodesic_kernel.cu:306-319)
    const R a = spin * M, cr = s.cam_r;
    const R Sig = cr * cr + a * a * ct * ct;
    const R Del = cr * cr - R(2) * M * cr + a * a;
    const R s2 = st * st;
    s.g_tt = -(R(1) - R(2) * M * cr / Sig);
    s.g_tp = R(-2) * M * cr * a * s2 / Sig;
    s.g_rr = Sig / Del;
    s.g_thth = Sig;
    s.g_pp = s2 * ((cr * cr + a * a) * (cr * cr + a * a) - a * a * Del * s2) / Sig;
    s.M = M;
    s.a = a;
    s.r_in = rin * M;
    s.r_out = rout * M;
    s.peak_T = R(base.disc_peak_T) * tscale;
    s.f_obs = R(1) - R(2) * M / cr;
    s.turbulence = R(base.disc_turbulence);
    s.time_sec = R(base.time_sec);
    s.w = R(soft_w);
    s.h = R(base.h_step);
    s.escape_r = R(base.escape_r);
    const R dd = M * M - a * a;
    s.cut = R(1.02) * (M + hsqrt(dd > R(0) ? dd : R(0)));
    s.width = base.width;
    s.height = base.height;
    s.flags = base.flags;
    s.max_steps = base.max_steps;
    return s;
}

// Camera ray and conserved quantities (geodesic_kernel.cu:188-223, 303-347).
template <class R>
SING_HD bool diff_init(const DiffScene<R>& sc, std::uint32_t gx, std::uint32_t gy, KY<R>& y, R& E, R& L) {
    const R fx = R(double(gx)) + R(0.5), fy = R(double(gy)) + R(0.5);
    const R un = (R(2) * fx / sc.width_f - R(1)) * sc.aspect * sc.tan_half_fov;
    const R vn = -(R(2) * fy / sc.height_f - R(1)) * sc.tan_half_fov;
    R d[3];
    for (int i = 0; i < 3; ++i) d[i] = sc.fwd[i] + (sc.right[i] * un + sc.up[i] * vn);
    const R len = hsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const R ik = R(1) / hmax(len, R(1e-20));
    for (int i = 0; i < 3; ++i) d[i] = d[i] * ik;
    const R ur = d[0] * sc.rhat[0] + d[1] * sc.rhat[1] + d[2] * sc.rhat[2];
    const R uth = (d[0] * sc.that[0] + d[1] * sc.that[1] + d[2] * sc.that[2]) / sc.cam_r;
    const R uph = (d[0] * sc.phat[0] + d[1] * sc.phat[1] + d[2] * sc.phat[2]) / sc.uphi_den;
    const R Bq = R(2) * sc.g_tp * uph;
    const R Cq = sc.g_rr * ur * ur + sc.g_thth * uth * uth + sc.g_pp * uph * uph;
    const R disc = Bq * Bq - R(4) * sc.g_tt * Cq;
    if (disc < R(0)) return false;
    const R ut = (-Bq - hsqrt(disc)) / (R(2) * sc.g_tt);
    E = -(sc.g_tt * ut + sc.g_tp * uph);
    L = sc.g_tp * ut + sc.g_pp * uph;
    y.r = sc.cam_r;
    y.th = sc.cam_theta;
    y.ph = sc.cam_phi;
    y.pr = sc.g_rr * ur;
    y.pt = sc.g_thth * uth;
    return true;
}

// Prograde ISCO radius / M of spin a/M in [0, 1) (Bardeen, Press & Teukolsky 1972).
template <class R> SING_HD R isco_prograde(R a) {
    const R third = R(1.0 / 3.0);
    const R z1 = R(1) + hpow(R(1) - a * a, third) * (hpow(R(1) + a, third) + hpow(R(1) - a, third));
    const R z2 = hsqrt(R(3) * a * a + z1 * z1);
    return R(3) + z2 - hsqrt((R(3) - z1) * (R(3) + z1 + R(2) * z2));
}

// Soft annulus opacity: 1 inside [r_in + w, r_out - w], linear to 0 at r_in and
// r_out.  It vanishes at r_in, where the emission jumps from 0 to the colour of
// the blackbody's 1000 K floor, so the composite stays continuous in r_cross.
template <class R> SING_HD R soft_alpha(R rc, R r_in, R r_out, R w) {
    const R lo = (rc - r_in) / w;
    const R hi = (r_out - rc) / w;
    return hclamp(hmin(lo, hi), R(0), R(1));
}

// Kerr disc emission at a crossing (geodesic_kernel.cu:77-121, 366-394).
template <class R>
SING_HD void diff_emit(R rc, R phi, R L, R a, R M, R r_in, R peak_T, R f_obs, R turb, R time_sec,
                       std::uint32_t flags, R* rgb) {
    rgb[0] = rgb[1] = rgb[2] = R(0);
    if (rc < r_in) return;
    const R ratio = r_in / rc;
    const R Tk = peak_T * (hpow(ratio, R(0.75)) * hmax(R(1) - hsqrt(ratio), R(0))) / R(0.186);
    if (!(Tk > R(0))) return;
    const R f_emit = R(1) - R(2) * M / rc;
    R g_grav = R(1), g_dop = R(1);
    if (flags & SING_FLAG_REDSHIFT_ON) g_grav = hsqrt(hmax(f_emit, R(1e-6)) / hmax(f_obs, R(1e-6)));
    if (flags & SING_FLAG_DOPPLER_ON) {
        R om = hsqrt(M / (rc * rc * rc));
        om = om / (R(1) + a * om);
        const R v = rc * om / hsqrt(hmax(f_emit, R(1e-6)));
        const R cos_psi = hclamp(L / rc, R(-1), R(1));
        g_dop = hsqrt(hmax(R(1) - v * v, R(0))) / hmax(R(1) - v * cos_psi, R(1e-4));
    }
    const R g = g_grav * g_dop;
    // blackbody (Tanner-Helland in centi-kelvin)
    const R T = hclamp(g * Tk, R(1000), R(40000)) * R(0.01);
    R cr, cg, cb;
    if (T <= R(66)) {
        cr = R(1);
        cg = hclamp(R(0.39008157876901960784) * hlog(T) - R(0.63184144378862745098), R(0), R(1));
    } else {
        const R t1 = T - R(60);
        cr = hclamp(R(1.29293618606274509804) * hpow(t1, R(-0.1332047592)), R(0), R(1));
        cg = hclamp(R(1.12989086089529411765) * hpow(t1, R(-0.0755148492)), R(0), R(1));
    }
    if (T >= R(66)) cb = R(1);
    else if (T <= R(19)) cb = R(0);
    else cb = hclamp(R(0.54320678911019607843) * hlog(T - R(10)) - R(1.19625408914), R(0), R(1));
    const R q = r_in / rc;
    const R flux = hclamp(q * q * q, R(0.05), R(1.5));
    // bands (geodesic_kernel.cu:109-121)
    const R x = rc / r_in;
    const R rho = hlog(hmax(x, R(0.1)));
    const R shear = R(4) / hmax(hpow(x, R(1.5)), R(0.25));
    const R omega = hpow(hmax(x, R(0.1)), R(-1.5));
    const R pa = phi - R(0.8) * time_sec * omega;
    const R b1 = R(0.5) + R(0.5) * hcos(R(10) * rho + R(3) * pa + shear);
    const R b2 = R(0.5) + R(0.5) * hcos(R(22) * rho - R(1.7) * pa);
    const R combined = R(0.8) + turb * (b1 * b1 * b2 - R(0.8));
    const R bands = hclamp(R(0.65) + R(0.55) * combined, R(0), R(1.3));
    const R g2 = g * g;
    const R wgt = flux * bands * g2 * g2;
    rgb[0] = cr * wgt;
    rgb[1] = cg * wgt;
    rgb[2] = cb * wgt;
}

}  // namespace sing
