// oracle/ref_exports.cpp — exports the reference's own host-compilable physics
// headers (shared_shader/*.h, core/include/physics/{kerr,disc,redshift}.hpp) as extern "C"
// functions, so the oracle's restatement can be checked against the real
// reference arithmetic.  Compiled ONLY from /root/reference by oracle/Makefile
// into oracle/_ref/libref.so; test infrastructure, never part of the product.
// No reference source is copied here: the headers are included from where
// they lie.

#include <cstddef>
#include <cstdint>

#include "disc_intersection.h"
#include "geodesic_math.h"
#include "kerr_hamilton.h"
#include "kerr_math.h"
#include "physics/disc.hpp"
#include "physics/kerr.hpp"
#include "physics/redshift.hpp"
#include "uniforms.h"

extern "C" {

int ref_sizeof_uniforms() { return int(sizeof(Uniforms)); }

// Byte offsets of the Uniforms fields in declaration order (36 fields, the
// vec3 members count once each).
int ref_uniform_offsets(int* out) {
    int i = 0;
    out[i++] = int(offsetof(Uniforms, cam_pos));
    out[i++] = int(offsetof(Uniforms, cam_fwd));
    out[i++] = int(offsetof(Uniforms, cam_right));
    out[i++] = int(offsetof(Uniforms, cam_up));
    out[i++] = int(offsetof(Uniforms, mass_M));
    out[i++] = int(offsetof(Uniforms, spin_a));
    out[i++] = int(offsetof(Uniforms, rs));
    out[i++] = int(offsetof(Uniforms, tan_half_fov));
    out[i++] = int(offsetof(Uniforms, disc_r_inner));
    out[i++] = int(offsetof(Uniforms, disc_r_outer));
    out[i++] = int(offsetof(Uniforms, disc_peak_T));
    out[i++] = int(offsetof(Uniforms, aspect));
    out[i++] = int(offsetof(Uniforms, h_step));
    out[i++] = int(offsetof(Uniforms, escape_r));
    out[i++] = int(offsetof(Uniforms, horizon_cut));
    out[i++] = int(offsetof(Uniforms, time_sec));
    out[i++] = int(offsetof(Uniforms, width));
    out[i++] = int(offsetof(Uniforms, height));
    out[i++] = int(offsetof(Uniforms, metric_type));
    out[i++] = int(offsetof(Uniforms, flags));
    out[i++] = int(offsetof(Uniforms, max_steps));
    out[i++] = int(offsetof(Uniforms, frame_index));
    out[i++] = int(offsetof(Uniforms, supersample));
    out[i++] = int(offsetof(Uniforms, pad_b));
    out[i++] = int(offsetof(Uniforms, exposure));
    out[i++] = int(offsetof(Uniforms, bloom_threshold));
    out[i++] = int(offsetof(Uniforms, bloom_strength));
    out[i++] = int(offsetof(Uniforms, disc_turbulence));
    return i;
}

static KerrConserved kc(const float* c) {
    KerrConserved k;
    k.E = c[0];
    k.L_z = c[1];
    k.Q = c[2];
    k.a = c[3];
    k.M = c[4];
    return k;
}

void ref_kerr_rhs(const float* s, const float* c, float* d) {
    KerrHamState y{s[0], s[1], s[2], s[3], s[4], s[5]};
    const KerrHamState r = kerr_ham_rhs(y, kc(c));
    d[0] = r.t; d[1] = r.r; d[2] = r.theta; d[3] = r.phi; d[4] = r.p_r; d[5] = r.p_theta;
}
float ref_kerr_two_H(const float* s, const float* c) { return kerr_ham_two_H(s[1], s[2], s[4], s[5], kc(c)); }
void ref_kerr_rk4(const float* s, float h, const float* c, int n, float* traj) {
    KerrHamState y{s[0], s[1], s[2], s[3], s[4], s[5]};
    const KerrConserved k = kc(c);
    for (int i = 0; i <= n; ++i) {
        float* o = traj + 6 * i;
        o[0] = y.t; o[1] = y.r; o[2] = y.theta; o[3] = y.phi; o[4] = y.p_r; o[5] = y.p_theta;
        if (i < n) y = kerr_ham_rk4_step(y, h, k);
    }
}
float ref_kerr_Theta(float theta, const float* c) { return kerr_Theta(theta, kc(c)); }
void ref_schw_rhs(const float* s, float rs, float* d) {
    State y{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
    const State r = geodesic_rhs_schwarzschild(y, rs);
    d[0] = r.t; d[1] = r.r; d[2] = r.theta; d[3] = r.phi; d[4] = r.ut; d[5] = r.ur; d[6] = r.utheta; d[7] = r.uphi;
}
void ref_schw_rk4(const float* s, float h, float rs, int n, float* traj) {
    State y{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
    for (int i = 0; i <= n; ++i) {
        float* o = traj + 8 * i;
        o[0] = y.t; o[1] = y.r; o[2] = y.theta; o[3] = y.phi; o[4] = y.ut; o[5] = y.ur; o[6] = y.utheta; o[7] = y.uphi;
        if (i < n) y = rk4_step(y, h, rs);
    }
}
void ref_crossing(float a, float b, int* crossed, float* frac) {
    const DiscCrossing x = detect_equatorial_crossing(a, b);
    *crossed = x.crossed ? 1 : 0;
    *frac = x.frac;
}
float ref_lerp(float a, float b, float f) { return lerp_scalar(a, b, f); }
int ref_in_annulus(float r, float a, float b) { return in_disc_annulus(r, a, b) ? 1 : 0; }
float ref_kerr_closed_form(int which, float M, float a) {
    using namespace singularity::physics;
    switch (which) {
        case 0: return kerr_outer_horizon(M, a);
        case 1: return kerr_inner_horizon(M, a);
        case 2: return kerr_isco_prograde(M, a);
        case 3: return kerr_isco_retrograde(M, a);
        case 4: return kerr_photon_sphere_prograde(M, a);
        case 5: return kerr_photon_sphere_retrograde(M, a);
    }
    return 0.0f;
}

// core/include/physics/disc.hpp:122-155: the host twin of the kernel's
// blackbody (no [1000, 40000] K clamp, constants divided by 255 at run time).
void ref_blackbody_th(float T, float* rgb) {
    const singularity::physics::sRGBColor c = singularity::physics::blackbody_srgb_tanner_helland(T);
    rgb[0] = c.r; rgb[1] = c.g; rgb[2] = c.b;
}
// core/include/physics/redshift.hpp:44-112
float ref_grav_redshift(float rs, float r_emit, float r_obs) {
    return singularity::physics::schwarzschild_grav_redshift(rs, r_emit, r_obs);
}
float ref_doppler_factor(float rs, float r_emit, float omega, float cos_psi) {
    return singularity::physics::schwarzschild_doppler_factor(rs, r_emit, omega, cos_psi);
}
float ref_keplerian_omega(float M, float r) { return singularity::physics::schwarzschild_keplerian_omega(M, r); }
float ref_intensity_scaling(float g) { return singularity::physics::intensity_scaling(g); }

}  // extern "C"
