// oracle/oracle_capi.cpp — extern "C" surface of the oracle for ctypes.
// Test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
// cpu_baseline leg); never linked into the product.

#include <cstdio>

#include "oracle.hpp"

using namespace orc;

namespace {

struct OrcExtras {  // mirrors orc::Extras in double; all zero = reference semantics
    double soft_edge;
    double band_phase1;
    double band_phase2;
    double adapt_tol;
    double peak_T_scale;  // 0 is read as 1
    double jitter_x;      // added to every sample offset
    double jitter_y;
    double grid_sampler;  // non-zero: the realtime backends' grid sampler
    double reproject;     // non-zero: Kerr null-constraint re-projection after every fixed step
    double cross_refine;  // non-zero: Kerr fixed-step crossing root from the FP64 dense output
};

template <class T> Uni<T> load_uni(const void* raw, const OrcExtras* ex) {
    UniformsRaw u;
    std::memcpy(&u, raw, sizeof(u));
    Uni<T> o = uni_from_raw<T>(u);
    if (ex) {
        o.ext.soft_edge = T(ex->soft_edge);
        o.ext.band_phase1 = T(ex->band_phase1);
        o.ext.band_phase2 = T(ex->band_phase2);
        o.ext.adapt_tol = T(ex->adapt_tol);
        o.ext.peak_T_scale = T(ex->peak_T_scale == 0.0 ? 1.0 : ex->peak_T_scale);
        o.ext.jitter_x = T(ex->jitter_x);
        o.ext.jitter_y = T(ex->jitter_y);
        o.ext.grid = ex->grid_sampler != 0.0;
        o.ext.reproject = ex->reproject != 0.0;
        o.ext.cross_refine = ex->cross_refine != 0.0;
    }
    return o;
}

template <class T>
void render_impl(const void* raw, const OrcExtras* ex, const std::int64_t* px, std::int64_t n, int threads,
                 double* hdr, double* ldr, std::uint8_t* rgba, std::uint8_t* cls, std::uint32_t* steps,
                 double* drift) {
    const Uni<T> u = load_uni<T>(raw, ex);
    const CamFrame<T> f = camera_frame(u);
    std::uint32_t spp = u.supersample < 1u ? 1u : (u.supersample > 1024u ? 1024u : u.supersample);
    std::vector<float> hal(2 * spp);
    halton_table(spp, hal.data());
    const std::int64_t W = u.width;
    parallel_for(n, threads, [&](std::int64_t i) {
        const std::int64_t p = px ? px[i] : i;
        const std::uint32_t gx = std::uint32_t(p % W), gy = std::uint32_t(p / W);
        const PixelOut<T> o = render_pixel(u, f, gx, gy, hal.data());
        for (int k = 0; k < 3; ++k) {
            if (hdr) hdr[3 * i + k] = o.hdr[k];
            if (ldr) ldr[3 * i + k] = o.ldr[k];
            if (drift) drift[3 * i + k] = o.inv_drift[k];
        }
        if (rgba) std::memcpy(rgba + 4 * i, o.rgba, 4);
        if (cls) cls[i] = o.cls;
        if (steps) steps[i] = o.steps;
    });
}

// Whole image rows, as the reference's CPU renderer schedules them
// (cli/main.cpp:739-740,861-872: hardware_concurrency() workers pulling the next
// row from an atomic counter): `rows` lists the rows to render, rgba receives
// n_rows * W * 4 bytes (row-major, in list order).  Returns the RK4 steps.
template <class T>
std::uint64_t rows_impl(const void* raw, const std::int32_t* rows, std::int64_t n_rows, int threads,
                        std::uint8_t* rgba) {
    const Uni<T> u = load_uni<T>(raw, nullptr);
    const CamFrame<T> f = camera_frame(u);
    std::uint32_t spp = u.supersample < 1u ? 1u : (u.supersample > 1024u ? 1024u : u.supersample);
    std::vector<float> hal(2 * spp);
    halton_table(spp, hal.data());
    const std::uint32_t W = u.width;
    std::atomic<std::uint64_t> total{0};
    parallel_for(n_rows, threads, [&](std::int64_t i) {
        std::uint64_t steps = 0;
        for (std::uint32_t x = 0; x < W; ++x) {
            const PixelOut<T> o = render_pixel(u, f, x, std::uint32_t(rows[i]), hal.data());
            if (rgba) std::memcpy(rgba + 4 * (std::size_t(i) * W + x), o.rgba, 4);
            steps += o.steps;
        }
        total.fetch_add(steps, std::memory_order_relaxed);
    });
    return total.load();
}

template <class T> void ray_impl(const void* raw, const OrcExtras* ex, std::uint32_t gx, std::uint32_t gy, double ox,
                                 double oy, double* out) {
    const Uni<T> u = load_uni<T>(raw, ex);
    const CamFrame<T> f = camera_frame(u);
    const RayOut<T> r = trace_ray(u, f, gx, gy, T(ox), T(oy));
    out[0] = val(r.rgb.x);
    out[1] = val(r.rgb.y);
    out[2] = val(r.rgb.z);
    out[3] = r.cls;
    out[4] = r.steps;
    out[5] = val(r.r_cross);
    for (int k = 0; k < 3; ++k) {
        out[6 + k] = r.inv0[k];
        out[9 + k] = r.inv1[k];
    }
    out[12] = r.E;
    out[13] = r.Lz;
    out[14] = r.r_min;
    out[15] = r.th_min;
    out[16] = r.esc_theta;
    out[17] = r.esc_phi;
    out[18] = r.th_max;
}

// Scene parameters of the differentiable renderer -> Uniforms in T.
// Orbital camera: cuda_cli/main.cpp:68-109; packing: cuda_backend.cu:45-101.
template <class T> struct Params {
    T spin, incl, r_in, r_out, tscale;
};
template <class T>
Uni<T> make_uni(const void* base_raw, const OrcExtras* ex, const Params<T>& p, double cam_dist, double az,
                double fov_deg) {
    Uni<T> u = load_uni<T>(base_raw, ex);
    const T half_pi = T(1.5707963267948966192313216916397514);
    const T elev = half_pi - p.incl;
    const T d = T(cam_dist);
    const T ce = xcos(elev), se = xsin(elev), ca = T(std::cos(az)), sa = T(std::sin(az));
    V3<T> pos = v3(d * ce * ca, d * ce * sa, d * se);
    const T inv = T(1.0) / d;
    const V3<T> fwd = v3(-pos.x * inv, -pos.y * inv, -pos.z * inv);
    V3<T> rt = v3(fwd.y, -fwd.x, T(0.0));
    const T rl = xsqrt(rt.x * rt.x + rt.y * rt.y + rt.z * rt.z);
    const T ir = (rl > T(1e-6)) ? T(1.0) / rl : T(1.0);
    rt = rt * ir;
    const V3<T> up = v3(fwd.y * rt.z - fwd.z * rt.y, fwd.z * rt.x - fwd.x * rt.z, fwd.x * rt.y - fwd.y * rt.x);
    u.cam_pos = pos;
    u.cam_right = rt;
    u.cam_up = up;
    u.cam_fwd = fwd;  // pack negates basis row 2, which is -fwd
    u.tan_half_fov = xtan(T(0.5) * T(fov_deg * (3.14159265358979323846 / 180.0)));
    const T M = u.mass_M;
    u.spin_a = p.spin * M;
    u.metric_type = (val(p.spin) > 0.0) ? METRIC_KERR : METRIC_SCHW;
    u.disc_r_inner = p.r_in * M;
    u.disc_r_outer = p.r_out * M;
    u.ext.peak_T_scale = p.tscale;
    return u;
}

}  // namespace

extern "C" {

int orc_sizeof_uniforms() { return int(sizeof(UniformsRaw)); }
float orc_halton(std::uint32_t i, std::uint32_t base) { return halton(i, base); }
void orc_halton_table(std::uint32_t spp, float* out) { halton_table(spp, out); }

int orc_render(const void* uniforms, int precision, const OrcExtras* ex, const std::int64_t* px, std::int64_t n,
               int threads, double* hdr, double* ldr, std::uint8_t* rgba, std::uint8_t* cls, std::uint32_t* steps,
               double* drift) {
    if (precision == 32) render_impl<float>(uniforms, ex, px, n, threads, hdr, ldr, rgba, cls, steps, drift);
    else if (precision == 64) render_impl<double>(uniforms, ex, px, n, threads, hdr, ldr, rgba, cls, steps, drift);
    else return -1;
    return 0;
}

long long orc_render_rows(const void* uniforms, int precision, const std::int32_t* rows, std::int64_t n_rows,
                          int threads, std::uint8_t* rgba) {
    if (precision == 32) return (long long)rows_impl<float>(uniforms, rows, n_rows, threads, rgba);
    if (precision == 64) return (long long)rows_impl<double>(uniforms, rows, n_rows, threads, rgba);
    return -1;
}

int orc_trace_ray(const void* uniforms, int precision, const OrcExtras* ex, std::uint32_t gx, std::uint32_t gy,
                  double ox, double oy, double* out19) {
    if (precision == 32) ray_impl<float>(uniforms, ex, gx, gy, ox, oy, out19);
    else if (precision == 64) ray_impl<double>(uniforms, ex, gx, gy, ox, oy, out19);
    else return -1;
    return 0;
}

// --- single functions (float and double) ---------------------------------
void orc_kerr_rhs_f32(const float* s, const float* c, float* d) {
    const KState<float> y{s[0], s[1], s[2], s[3], s[4], s[5]};
    const KConst<float> k{c[0], c[1], c[2], c[3], c[4]};
    const KState<float> r = kerr_rhs(y, k);
    d[0] = r.t; d[1] = r.r; d[2] = r.theta; d[3] = r.phi; d[4] = r.p_r; d[5] = r.p_theta;
}
void orc_kerr_rhs_f64(const double* s, const double* c, double* d) {
    const KState<double> y{s[0], s[1], s[2], s[3], s[4], s[5]};
    const KConst<double> k{c[0], c[1], c[2], c[3], c[4]};
    const KState<double> r = kerr_rhs(y, k);
    d[0] = r.t; d[1] = r.r; d[2] = r.theta; d[3] = r.phi; d[4] = r.p_r; d[5] = r.p_theta;
}
float orc_kerr_two_H_f32(const float* s, const float* c) {
    return kerr_two_H(s[1], s[2], s[4], s[5], KConst<float>{c[0], c[1], c[2], c[3], c[4]});
}
double orc_kerr_two_H_f64(const double* s, const double* c) {
    return kerr_two_H(s[1], s[2], s[4], s[5], KConst<double>{c[0], c[1], c[2], c[3], c[4]});
}
// Null-constraint re-projection (kerr_reproject) of s[6] = (t, r, theta, phi,
// p_r, p_theta) with c[5] = (E, L_z, Q, a, M); out[6] = the projected state.
void orc_kerr_reproject_f32(const float* s, const float* c, float* out) {
    const KState<float> o = kerr_reproject(KState<float>{s[0], s[1], s[2], s[3], s[4], s[5]},
                                           KConst<float>{c[0], c[1], c[2], c[3], c[4]});
    out[0] = o.t; out[1] = o.r; out[2] = o.theta; out[3] = o.phi; out[4] = o.p_r; out[5] = o.p_theta;
}
void orc_kerr_reproject_f64(const double* s, const double* c, double* out) {
    const KState<double> o = kerr_reproject(KState<double>{s[0], s[1], s[2], s[3], s[4], s[5]},
                                            KConst<double>{c[0], c[1], c[2], c[3], c[4]});
    out[0] = o.t; out[1] = o.r; out[2] = o.theta; out[3] = o.phi; out[4] = o.p_r; out[5] = o.p_theta;
}
// FP64-refined equator crossing (kerr_refined_crossing) of the step y0 -> y1 of
// size h; out[2] = (r, phi) at the root.
void orc_kerr_refined_crossing_f32(const float* y0, const float* y1, float h, const float* c, double* out) {
    const DenseCross<float> d = kerr_refined_crossing(KState<float>{y0[0], y0[1], y0[2], y0[3], y0[4], y0[5]},
                                                      KState<float>{y1[0], y1[1], y1[2], y1[3], y1[4], y1[5]}, h,
                                                      KConst<float>{c[0], c[1], c[2], c[3], c[4]});
    out[0] = d.r;
    out[1] = d.phi;
}
void orc_kerr_refined_crossing_f64(const double* y0, const double* y1, double h, const double* c, double* out) {
    const DenseCross<double> d = kerr_refined_crossing(KState<double>{y0[0], y0[1], y0[2], y0[3], y0[4], y0[5]},
                                                       KState<double>{y1[0], y1[1], y1[2], y1[3], y1[4], y1[5]}, h,
                                                       KConst<double>{c[0], c[1], c[2], c[3], c[4]});
    out[0] = d.r;
    out[1] = d.phi;
}
// n steps of the damped RK4; traj receives 6*(n+1) values (state after each step, incl. start).
void orc_kerr_rk4_f32(const float* s, float h, const float* c, int n, float* traj) {
    KState<float> y{s[0], s[1], s[2], s[3], s[4], s[5]};
    const KConst<float> k{c[0], c[1], c[2], c[3], c[4]};
    for (int i = 0; i <= n; ++i) {
        float* o = traj + 6 * i;
        o[0] = y.t; o[1] = y.r; o[2] = y.theta; o[3] = y.phi; o[4] = y.p_r; o[5] = y.p_theta;
        if (i < n) y = kerr_rk4(y, h, k);
    }
}
void orc_kerr_rk4_f64(const double* s, double h, const double* c, int n, double* traj) {
    KState<double> y{s[0], s[1], s[2], s[3], s[4], s[5]};
    const KConst<double> k{c[0], c[1], c[2], c[3], c[4]};
    for (int i = 0; i <= n; ++i) {
        double* o = traj + 6 * i;
        o[0] = y.t; o[1] = y.r; o[2] = y.theta; o[3] = y.phi; o[4] = y.p_r; o[5] = y.p_theta;
        if (i < n) y = kerr_rk4(y, h, k);
    }
}
void orc_schw_rhs_f32(const float* s, float rs, float* d) {
    const SState<float> y{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
    const SState<float> r = schw_rhs(y, rs);
    d[0] = r.t; d[1] = r.r; d[2] = r.theta; d[3] = r.phi; d[4] = r.ut; d[5] = r.ur; d[6] = r.utheta; d[7] = r.uphi;
}
void orc_schw_rhs_f64(const double* s, double rs, double* d) {
    const SState<double> y{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
    const SState<double> r = schw_rhs(y, rs);
    d[0] = r.t; d[1] = r.r; d[2] = r.theta; d[3] = r.phi; d[4] = r.ut; d[5] = r.ur; d[6] = r.utheta; d[7] = r.uphi;
}
void orc_schw_rk4_f32(const float* s, float h, float rs, int n, float* traj) {
    SState<float> y{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
    for (int i = 0; i <= n; ++i) {
        float* o = traj + 8 * i;
        o[0] = y.t; o[1] = y.r; o[2] = y.theta; o[3] = y.phi; o[4] = y.ut; o[5] = y.ur; o[6] = y.utheta; o[7] = y.uphi;
        if (i < n) y = schw_rk4(y, h, rs);
    }
}
void orc_schw_rk4_f64(const double* s, double h, double rs, int n, double* traj) {
    SState<double> y{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
    for (int i = 0; i <= n; ++i) {
        double* o = traj + 8 * i;
        o[0] = y.t; o[1] = y.r; o[2] = y.theta; o[3] = y.phi; o[4] = y.ut; o[5] = y.ur; o[6] = y.utheta; o[7] = y.uphi;
        if (i < n) y = schw_rk4(y, h, rs);
    }
}
void orc_crossing_f32(float a, float b, int* crossed, float* frac) {
    const Crossing<float> x = detect_crossing(a, b);
    *crossed = x.crossed;
    *frac = x.frac;
}
void orc_blackbody_f32(float T, float* rgb) {
    const V3<float> c = blackbody_srgb(T);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
void orc_blackbody_f64(double T, double* rgb) {
    const V3<double> c = blackbody_srgb(T);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
float orc_disc_temperature_f32(float r, float rin, float peak) { return disc_temperature(r, rin, peak); }
double orc_disc_temperature_f64(double r, double rin, double peak) { return disc_temperature(r, rin, peak); }
float orc_band_mask_f32(float r, float phi, float rin, float turb, float t, float ph1, float ph2) {
    return disc_band_mask(r, phi, rin, turb, t, ph1, ph2);
}
void orc_schw_shift_f32(float rs, float M, float r, float uphi, float f_cam, unsigned flags, float* out) {
    const Shift<float> sh = schw_shift(rs, M, r, uphi, f_cam, flags);
    out[0] = sh.g_grav; out[1] = sh.g_dop; out[2] = sh.omega;
}
void orc_kerr_shift_f32(float M, float a, float r, float cam_r, float L_z, unsigned flags, float* out) {
    const Shift<float> sh = kerr_shift(M, a, r, cam_r, L_z, flags);
    out[0] = sh.g_grav; out[1] = sh.g_dop; out[2] = sh.omega;
}
void orc_kerr_shift_f64(double M, double a, double r, double cam_r, double L_z, unsigned flags, double* out) {
    const Shift<double> sh = kerr_shift(M, a, r, cam_r, L_z, flags);
    out[0] = sh.g_grav; out[1] = sh.g_dop; out[2] = sh.omega;
}
float orc_disc_flux_f32(float r_in, float r) { return disc_flux(r_in, r); }
// Disc emission of the uniform block's scene at a crossing, Kerr or
// Schwarzschild by metric_type: x = cam_r (Kerr) or f_cam (Schwarzschild),
// y = L_z (Kerr) or u^phi at the crossing (Schwarzschild).
void orc_disc_emit_f32(const void* uniforms, float r, float phi, float x, float y, float Tk, float* rgb) {
    const Uni<float> u = load_uni<float>(uniforms, nullptr);
    const V3<float> c = u.metric_type == 1u ? kerr_disc_emit(u, r, phi, x, y, Tk) : schw_disc_emit(u, r, y, phi, x, Tk);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
void orc_starfield_f32(float x, float y, float z, float* rgb) {
    const V3<float> c = starfield_color(v3(x, y, z));
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
std::uint32_t orc_mix32(std::uint32_t a, std::uint32_t b) { return mix32(a, b); }
void orc_aces_f32(const float* in, float* out) {
    const V3<float> c = aces(v3(in[0], in[1], in[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
float orc_linear_to_srgb_f32(float c) { return linear_to_srgb(c); }
std::uint8_t orc_to_u8_f32(float c) { return to_u8(c); }
void orc_camera_ray_f32(const void* uniforms, std::uint32_t gx, std::uint32_t gy, float ox, float oy, float* out) {
    const Uni<float> u = load_uni<float>(uniforms, nullptr);
    const CamFrame<float> f = camera_frame(u);
    const CameraRay<float> r = camera_ray(u, f, gx, gy, ox, oy);
    out[0] = r.dir.x; out[1] = r.dir.y; out[2] = r.dir.z;
    out[3] = r.ur; out[4] = r.utheta; out[5] = r.uphi;
    out[6] = f.cam_r; out[7] = f.cam_theta; out[8] = f.cam_phi; out[9] = f.sin_th;
}
double orc_kerr_closed_form(int which, double M, double a) {
    switch (which) {
        case 0: return kerr_r_plus(M, a);
        case 1: return kerr_r_minus(M, a);
        case 2: return kerr_isco(M, a, true);
        case 3: return kerr_isco(M, a, false);
        case 4: return kerr_photon_pro(M, a);
        case 5: return kerr_photon_ret(M, a);
    }
    return 0.0;
}
float orc_kerr_closed_form_f32(int which, float M, float a) {
    switch (which) {
        case 0: return kerr_r_plus(M, a);
        case 1: return kerr_r_minus(M, a);
        case 2: return kerr_isco(M, a, true);
        case 3: return kerr_isco(M, a, false);
        case 4: return kerr_photon_pro(M, a);
        case 5: return kerr_photon_ret(M, a);
    }
    return 0.0f;
}

// --- post passes ------------------------------------------------------------
void orc_to_half(float* x, std::int64_t n) {
    for (std::int64_t i = 0; i < n; ++i) x[i] = to_half_and_back(x[i]);
}
void orc_bloom_extract(const float* hdr, int W, int H, float thr, float* out, int half_store) {
    bloom_extract(hdr, W, H, thr, out, half_store != 0);
}
void orc_bloom_blur(const float* in, int Wb, int Hb, int horizontal, float* out, int half_store) {
    bloom_blur(in, Wb, Hb, horizontal != 0, out, half_store != 0);
}
void orc_bloom_composite(const float* hdr, const float* bloom, int W, int H, float strength, float exposure,
                         std::uint8_t* rgba) {
    bloom_composite(hdr, bloom, W, H, strength, exposure, rgba);
}

// --- algorithmic operation counts per RK4 step ------------------------------
// counts[0..5] = add, mul, div, sqrt, trig, other for one step at a generic state.
void orc_count_kerr_step(double* counts) {
    tally() = OpTally{};
    const KState<OpCount> y{0.0, 8.0, 1.2, 0.3, 0.4, 1.1};
    const KConst<OpCount> c{1.0, 2.5, 6.0, 0.94, 1.0};
    (void)kerr_rk4(y, OpCount(0.4), c);
    const OpTally t = tally();
    counts[0] = double(t.add); counts[1] = double(t.mul); counts[2] = double(t.div);
    counts[3] = double(t.sqrt); counts[4] = double(t.trig); counts[5] = double(t.other);
}
void orc_count_schw_step(double* counts) {
    tally() = OpTally{};
    const SState<OpCount> y{0.0, 8.0, 1.2, 0.3, 1.2, -0.4, 0.01, 0.02};
    (void)schw_rk4(y, OpCount(0.4), OpCount(2.0));
    const OpTally t = tally();
    counts[0] = double(t.add); counts[1] = double(t.mul); counts[2] = double(t.div);
    counts[3] = double(t.sqrt); counts[4] = double(t.trig); counts[5] = double(t.other);
}

// --- differentiable scene: uniforms from parameters, and the dual-number gradient
// params[5] = spin a/M, inclination (rad), r_in/M, r_out/M, peak-T scale.
void orc_make_uniforms(const void* base, const double* params, double cam_dist, double az, double fov_deg,
                       void* out_raw) {
    const Params<double> p{params[0], params[1], params[2], params[3], params[4]};
    const Uni<double> u = make_uni<double>(base, nullptr, p, cam_dist, az, fov_deg);
    UniformsRaw o;
    std::memcpy(&o, base, sizeof(o));
    o.cam_pos = V4{float(u.cam_pos.x), float(u.cam_pos.y), float(u.cam_pos.z), 0.0f};
    o.cam_fwd = V4{float(u.cam_fwd.x), float(u.cam_fwd.y), float(u.cam_fwd.z), 0.0f};
    o.cam_right = V4{float(u.cam_right.x), float(u.cam_right.y), float(u.cam_right.z), 0.0f};
    o.cam_up = V4{float(u.cam_up.x), float(u.cam_up.y), float(u.cam_up.z), 0.0f};
    o.tan_half_fov = float(u.tan_half_fov);
    o.spin_a = float(u.spin_a);
    o.metric_type = u.metric_type;
    o.disc_r_inner = float(u.disc_r_inner);
    o.disc_r_outer = float(u.disc_r_outer);
    std::memcpy(out_raw, &o, sizeof(o));
}

// Loss = sum over listed pixels and RGB of (hdr - target)^2, hdr in double with
// forward-mode derivatives wrt the 5 parameters.  Returns the loss; grad[5].
double orc_loss_grad_dual(const void* base, const OrcExtras* ex, const double* params, double cam_dist, double az,
                          double fov_deg, const std::int64_t* px, std::int64_t n, const double* target, int threads,
                          double* grad, double* hdr_out) {
    using D = Dual<5>;
    const Params<D> p{D::var(params[0], 0), D::var(params[1], 1), D::var(params[2], 2), D::var(params[3], 3),
                      D::var(params[4], 4)};
    const Uni<D> u = make_uni<D>(base, ex, p, cam_dist, az, fov_deg);
    const CamFrame<D> f = camera_frame(u);
    std::vector<double> part(std::size_t(n) * 6, 0.0), lossv(std::size_t(n), 0.0);
    const std::int64_t W = u.width;
    parallel_for(n, threads, [&](std::int64_t i) {
        const std::int64_t pi = px ? px[i] : i;
        const std::uint32_t gx = std::uint32_t(pi % W), gy = std::uint32_t(pi / W);
        const RayOut<D> r = trace_ray(u, f, gx, gy, D(0.5), D(0.5));
        const D c[3] = {r.rgb.x, r.rgb.y, r.rgb.z};
        double l = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double diff = c[k].v - target[3 * i + k];
            l += diff * diff;
            for (int j = 0; j < 5; ++j) part[6 * i + j] += 2.0 * diff * c[k].d[j];
            if (hdr_out) hdr_out[3 * i + k] = c[k].v;
        }
        lossv[i] = l;
    });
    double loss = 0.0;
    for (int j = 0; j < 5; ++j) grad[j] = 0.0;
    for (std::int64_t i = 0; i < n; ++i) {
        loss += lossv[i];
        for (int j = 0; j < 5; ++j) grad[j] += part[6 * i + j];
    }
    return loss;
}

// Loss only (double), for central finite differences of the same model.
double orc_loss_f64(const void* base, const OrcExtras* ex, const double* params, double cam_dist, double az,
                    double fov_deg, const std::int64_t* px, std::int64_t n, const double* target, int threads,
                    double* hdr_out) {
    const Params<double> p{params[0], params[1], params[2], params[3], params[4]};
    const Uni<double> u = make_uni<double>(base, ex, p, cam_dist, az, fov_deg);
    const CamFrame<double> f = camera_frame(u);
    std::vector<double> lossv(std::size_t(n), 0.0);
    const std::int64_t W = u.width;
    parallel_for(n, threads, [&](std::int64_t i) {
        const std::int64_t pi = px ? px[i] : i;
        const RayOut<double> r = trace_ray(u, f, std::uint32_t(pi % W), std::uint32_t(pi / W), 0.5, 0.5);
        const double c[3] = {r.rgb.x, r.rgb.y, r.rgb.z};
        double l = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double diff = c[k] - (target ? target[3 * i + k] : 0.0);
            l += diff * diff;
            if (hdr_out) hdr_out[3 * i + k] = c[k];
        }
        lossv[i] = l;
    });
    double loss = 0.0;
    for (std::int64_t i = 0; i < n; ++i) loss += lossv[i];
    return loss;
}

}  // extern "C"

extern "C" {
// Exact Jacobian (forward-mode dual numbers) of one damped Kerr RK4 step of the
// reference algorithm (kerr_hamilton.h:165-178) in double:
// J[i*8 + j] = d y'_i / d x_j, y = (r, theta, phi, p_r, p_theta), x = (y, E, L_z, a).
void orc_kerr_step_jacobian(const double* y, double E, double L, double a, double M, double h, double* J) {
    using D = Dual<8>;
    const KState<D> s{D(0.0), D::var(y[0], 0), D::var(y[1], 1), D::var(y[2], 2), D::var(y[3], 3), D::var(y[4], 4)};
    const KConst<D> c{D::var(E, 5), D::var(L, 6), D(0.0), D::var(a, 7), D(M)};
    const KState<D> o = kerr_rk4(s, D(h), c);
    const D* out[5] = {&o.r, &o.theta, &o.phi, &o.p_r, &o.p_theta};
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 8; ++j) J[i * 8 + j] = out[i]->d[j];
}
}

extern "C" {
// One error-controlled step (the oracle's kerr_step with adapt_tol = tol) in
// double.  y[5] = (r, theta, phi, p_r, p_theta); c[4] = (E, L_z, a, M).
// h_cur is the controller state (in/out).  out[5] = accepted state;
// returns the step size actually taken (h_cur at acceptance x pole scale).
double orc_kerr_adaptive_step(const double* y, const double* c, double h_step, double tol, double* h_cur,
                              double* out) {
    Uni<double> u{};
    u.h_step = h_step;
    u.ext.adapt_tol = tol;
    const KState<double> s{0.0, y[0], y[1], y[2], y[3], y[4]};
    const KConst<double> k{c[0], c[1], 0.0, c[2], c[3]};
    double h = *h_cur;
    const KState<double> o = kerr_step(s, k, u, h);
    // the step taken: replay the controller's halvings to find h at acceptance
    double hc = *h_cur;
    for (int attempt = 0; attempt < 12; ++attempt) {
        const double hh = hc * kerr_pole_scale(s.theta);
        const KState<double> half = kerr_rk4_raw(kerr_rk4_raw(s, 0.5 * hh, k), 0.5 * hh, k);
        if (half.r == o.r && half.theta == o.theta && half.phi == o.phi && half.p_r == o.p_r &&
            half.p_theta == o.p_theta)
            break;
        hc *= 0.5;
    }
    *h_cur = h;
    out[0] = o.r; out[1] = o.theta; out[2] = o.phi; out[3] = o.p_r; out[4] = o.p_theta;
    return hc * kerr_pole_scale(s.theta);
}
// Reference solution for the same step: n classical RK4 substeps of h/n.
void orc_kerr_rk4_substeps(const double* y, const double* c, double h, int n, double* out) {
    KState<double> s{0.0, y[0], y[1], y[2], y[3], y[4]};
    const KConst<double> k{c[0], c[1], 0.0, c[2], c[3]};
    for (int i = 0; i < n; ++i) s = kerr_rk4_raw(s, h / n, k);
    out[0] = s.r; out[1] = s.theta; out[2] = s.phi; out[3] = s.p_r; out[4] = s.p_theta;
}
}

extern "C" {
// The CPU driver's diagnostics (cli/main.cpp:1269-1621); precision 32 runs the
// reference's float arithmetic, 64 the same steps in double.
int orc_kerr_geometry(int precision, double M, double a, double* out11) {
    if (precision == 32) {
        float o[11];
        const int rc = kerr_geometry<float>(float(M), float(a), o);
        for (int i = 0; i < 11; ++i) out11[i] = o[i];
        return rc;
    }
    return kerr_geometry<double>(M, a, out11);
}
// rows: cap x 5 doubles (step, t, r, theta, phi); *n_rows = rows produced
// (may exceed cap: then only cap are written).  Returns 0 or 2 (argument error).
int orc_photon_orbit(int precision, double M, double a, double r_init, double orbits, double h, double* rows,
                     int cap, int* n_rows) {
    std::vector<OrbitRow> v;
    const int rc = precision == 32 ? photon_orbit<float>(float(M), float(a), float(r_init), float(orbits), float(h), v)
                                   : photon_orbit<double>(M, a, r_init, orbits, h, v);
    *n_rows = int(v.size());
    for (int i = 0; i < int(v.size()) && i < cap; ++i) {
        rows[5 * i + 0] = v[i].step;
        rows[5 * i + 1] = v[i].t;
        rows[5 * i + 2] = v[i].r;
        rows[5 * i + 3] = v[i].theta;
        rows[5 * i + 4] = v[i].phi;
    }
    return rc;
}
long long orc_bench_fan(int precision, int kerr, double spin, int n_rays, int max_steps, double b_min, double b_max,
                        std::uint32_t* per_ray) {
    if (precision == 32)
        return bench_fan<float>(kerr != 0, float(spin), n_rays, max_steps, float(b_min), float(b_max), per_ray);
    return bench_fan<double>(kerr != 0, spin, n_rays, max_steps, b_min, b_max, per_ray);
}
}
