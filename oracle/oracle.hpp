// oracle/oracle.hpp — CPU restatement of the reference's offline null-geodesic
// renderer, used ONLY as test infrastructure: by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
// (singularity_amd/, include/) never includes, links or calls it.
//
// Everything is templated on the scalar type T so the same discrete algorithm
// runs as
//   * float   — reference semantics (the CUDA path is all-FP32),
//   * double  — the FP64 oracle: same steps, clamps, termination order and
//               literals evaluated in double (pi/2 exact to double),
//   * Dual<N> — forward-mode dual numbers in double: exact derivatives of the
//               discrete algorithm (gradient oracle for the adjoint),
//   * OpCount — a double that counts its arithmetic (algorithmic flop counts).
//
// Each function cites the reference file:line whose arithmetic it follows.
// The kerr/schwarzschild RHS keep the reference's expression order so the
// float instantiation is bit-identical to the reference headers built with
// -ffp-contract=off (checked against oracle/_ref by tests).
//
// Every function here is pinned by a test (tests/test_oracle_vs_reference.py,
// tests/test_oracle_math.py, tests/test_oracle_shading.py,
// tests/test_golden_images.py): the physics bit-exact against the reference
// headers, the shading against the reference's host twins
// (core/include/physics/{disc,redshift}.hpp) and hand-derived known answers,
// the bloom passes by their closed forms (uniform blocks, impulses, the knee
// at the threshold, the weight sum); the options beyond the reference by
// convergence and closed forms (tests/test_oracle_adaptive.py: error-controlled
// steps; tests/test_oracle_refine.py: the refined crossing against straight
// lines in flat space, the null-constraint re-projection against the exact
// second-order residual of its Newton step).

#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Scalar shims.  Generic code calls x*() so float/double/Dual/OpCount all work.
// ---------------------------------------------------------------------------
inline float xsin(float x) { return std::sin(x); }
inline double xsin(double x) { return std::sin(x); }
inline float xcos(float x) { return std::cos(x); }
inline double xcos(double x) { return std::cos(x); }
inline float xtan(float x) { return std::tan(x); }
inline double xtan(double x) { return std::tan(x); }
inline float xsqrt(float x) { return std::sqrt(x); }
inline double xsqrt(double x) { return std::sqrt(x); }
inline float xlog(float x) { return std::log(x); }
inline double xlog(double x) { return std::log(x); }
inline float xpow(float x, float y) { return std::pow(x, y); }
inline double xpow(double x, double y) { return std::pow(x, y); }
inline float xacos(float x) { return std::acos(x); }
inline double xacos(double x) { return std::acos(x); }
inline float xatan2(float y, float x) { return std::atan2(y, x); }
inline double xatan2(double y, double x) { return std::atan2(y, x); }
inline float xcbrt(float x) { return std::cbrt(x); }
inline double xcbrt(double x) { return std::cbrt(x); }
inline float xfmax(float a, float b) { return std::fmax(a, b); }
inline double xfmax(double a, double b) { return std::fmax(a, b); }
inline float xfmin(float a, float b) { return std::fmin(a, b); }
inline double xfmin(double a, double b) { return std::fmin(a, b); }
inline bool xisfinite(float x) { return std::isfinite(x); }
inline bool xisfinite(double x) { return std::isfinite(x); }
inline double val(float x) { return x; }
inline double val(double x) { return x; }

// ---------------------------------------------------------------------------
// Forward-mode dual number over double with N tangent directions.
// ---------------------------------------------------------------------------
template <int N>
struct Dual {
    double v = 0.0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: literals and constants promote
    static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};
template <int N> inline double val(const Dual<N>& x) { return x.v; }
template <int N> inline Dual<N> chain(const Dual<N>& x, double f, double df) {
    Dual<N> r(f);
    for (int i = 0; i < N; ++i) r.d[i] = df * x.d[i];
    return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N> inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { return a = a + b; }
template <int N> inline Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { return a = a * b; }
template <int N> inline bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> inline bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> inline bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N> inline bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N> inline Dual<N> xsin(const Dual<N>& x) { return chain(x, std::sin(x.v), std::cos(x.v)); }
template <int N> inline Dual<N> xcos(const Dual<N>& x) { return chain(x, std::cos(x.v), -std::sin(x.v)); }
template <int N> inline Dual<N> xtan(const Dual<N>& x) {
    const double t = std::tan(x.v);
    return chain(x, t, 1.0 + t * t);
}
template <int N> inline Dual<N> xsqrt(const Dual<N>& x) {
    const double s = std::sqrt(x.v);
    return chain(x, s, s > 0.0 ? 0.5 / s : 0.0);
}
template <int N> inline Dual<N> xlog(const Dual<N>& x) { return chain(x, std::log(x.v), 1.0 / x.v); }
template <int N> inline Dual<N> xpow(const Dual<N>& x, const Dual<N>& y) {
    // y is a constant exponent everywhere on this path (0.75, 1.5, 3, ...).
    const double p = std::pow(x.v, y.v);
    return chain(x, p, x.v != 0.0 ? y.v * p / x.v : 0.0);
}
template <int N> inline Dual<N> xacos(const Dual<N>& x) {
    return chain(x, std::acos(x.v), -1.0 / std::sqrt(std::max(1e-300, 1.0 - x.v * x.v)));
}
template <int N> inline Dual<N> xatan2(const Dual<N>& y, const Dual<N>& x) {
    Dual<N> r(std::atan2(y.v, x.v));
    const double q = x.v * x.v + y.v * y.v;
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / q;
    return r;
}
template <int N> inline Dual<N> xcbrt(const Dual<N>& x) {
    const double c = std::cbrt(x.v);
    return chain(x, c, c != 0.0 ? 1.0 / (3.0 * c * c) : 0.0);
}
template <int N> inline Dual<N> xfmax(const Dual<N>& a, const Dual<N>& b) { return (a.v >= b.v || std::isnan(b.v)) ? a : b; }
template <int N> inline Dual<N> xfmin(const Dual<N>& a, const Dual<N>& b) { return (a.v <= b.v || std::isnan(b.v)) ? a : b; }
template <int N> inline bool xisfinite(const Dual<N>& x) { return std::isfinite(x.v); }

// ---------------------------------------------------------------------------
// Operation-counting scalar: counts the arithmetic the algorithm performs.
// ---------------------------------------------------------------------------
struct OpTally {
    std::uint64_t add = 0, mul = 0, div = 0, sqrt = 0, trig = 0, other = 0;
};
inline OpTally& tally() {
    static thread_local OpTally t;
    return t;
}
struct OpCount {
    double v = 0.0;
    OpCount() = default;
    OpCount(double x) : v(x) {}  // NOLINT
};
inline double val(const OpCount& x) { return x.v; }
inline OpCount operator+(OpCount a, OpCount b) { ++tally().add; return a.v + b.v; }
inline OpCount operator-(OpCount a, OpCount b) { ++tally().add; return a.v - b.v; }
inline OpCount operator-(OpCount a) { return -a.v; }
inline OpCount operator*(OpCount a, OpCount b) { ++tally().mul; return a.v * b.v; }
inline OpCount operator/(OpCount a, OpCount b) { ++tally().div; return a.v / b.v; }
inline bool operator<(OpCount a, OpCount b) { return a.v < b.v; }
inline bool operator>(OpCount a, OpCount b) { return a.v > b.v; }
inline bool operator<=(OpCount a, OpCount b) { return a.v <= b.v; }
inline bool operator>=(OpCount a, OpCount b) { return a.v >= b.v; }
inline OpCount xsin(OpCount x) { ++tally().trig; return std::sin(x.v); }
inline OpCount xcos(OpCount x) { ++tally().trig; return std::cos(x.v); }
inline OpCount xtan(OpCount x) { ++tally().trig; return std::tan(x.v); }
inline OpCount xsqrt(OpCount x) { ++tally().sqrt; return std::sqrt(x.v); }
inline OpCount xlog(OpCount x) { ++tally().other; return std::log(x.v); }
inline OpCount xpow(OpCount x, OpCount y) { ++tally().other; return std::pow(x.v, y.v); }
inline OpCount xacos(OpCount x) { ++tally().other; return std::acos(x.v); }
inline OpCount xatan2(OpCount y, OpCount x) { ++tally().other; return std::atan2(y.v, x.v); }
inline OpCount xcbrt(OpCount x) { ++tally().other; return std::cbrt(x.v); }
inline OpCount xfmax(OpCount a, OpCount b) { return std::fmax(a.v, b.v); }
inline OpCount xfmin(OpCount a, OpCount b) { return std::fmin(a.v, b.v); }
inline bool xisfinite(OpCount x) { return std::isfinite(x.v); }

template <class T> inline T vclamp(T x, T lo, T hi) { return xfmin(xfmax(x, lo), hi); }
template <class T> inline T vmix(T a, T b, T t) { return a + t * (b - a); }

// ---------------------------------------------------------------------------
// Uniforms: the 160-byte host<->kernel block (shared_shader/uniforms.h:54-95)
// and a T-valued working copy.
// ---------------------------------------------------------------------------
enum : std::uint32_t {
    FLAG_DOPPLER = 1u << 0,    // uniforms.h:41
    FLAG_REDSHIFT = 1u << 1,   // uniforms.h:42
    FLAG_DISC = 1u << 2,       // uniforms.h:43
    FLAG_STARFIELD = 1u << 3,  // uniforms.h:44
    FLAG_ADAPTIVE = 1u << 4,   // uniforms.h:48
};
enum : std::uint32_t { METRIC_SCHW = 0u, METRIC_KERR = 1u };  // uniforms.h:51-52

struct alignas(16) V4 {
    float x, y, z, w;
};
struct UniformsRaw {  // byte layout of shared_shader/uniforms.h:54-95
    V4 cam_pos, cam_fwd, cam_right, cam_up;
    float mass_M, spin_a, rs, tan_half_fov;
    float disc_r_inner, disc_r_outer, disc_peak_T, aspect;
    float h_step, escape_r, horizon_cut, time_sec;
    std::uint32_t width, height, metric_type, flags;
    std::uint32_t max_steps, frame_index, supersample, pad_b;
    float exposure, bloom_threshold, bloom_strength, disc_turbulence;
};
static_assert(sizeof(UniformsRaw) == 160, "Uniforms must be 160 bytes");

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T k) { return {a.x * k, a.y * k, a.z * k}; }
template <class T> inline T vdot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T vlength(V3<T> a) { return xsqrt(vdot(a, a)); }
// geodesic_kernel.cu:49-62 (vdiv multiplies by the reciprocal)
template <class T> inline V3<T> vnormalize(V3<T> a) {
    const T k = xfmax(vlength(a), T(1e-20));
    const T ik = T(1.0) / k;
    return a * ik;
}

// Extensions the build adds on top of the 160-B block.  Defaults reproduce the
// reference exactly.
template <class T> struct Extras {
    T soft_edge = T(0.0);   // annulus edge width (M) of the differentiable disc model; 0 = hard edges
    T band_phase1 = T(0.0); // seeded turbulence phase offsets (seed 0 -> 0, 0)
    T band_phase2 = T(0.0);
    T adapt_tol = T(0.0);   // >0: error-controlled RK4 (step doubling), tolerance per step
    T peak_T_scale = T(1.0);
    T jitter_x = T(0.0);    // added to every sample offset (boundary-band probes)
    T jitter_y = T(0.0);
    bool grid = false;      // realtime sampler: ss x ss grid, ss = clamp(spp, 1, 4), tone map of the mean
    bool reproject = false;     // Kerr: null-constraint re-projection after every fixed RK4 step
    bool cross_refine = false;  // Kerr fixed step: crossing root from the FP64 cubic dense output
};

template <class T> struct Uni {
    V3<T> cam_pos, cam_fwd, cam_right, cam_up;
    T mass_M, spin_a, rs, tan_half_fov, disc_r_inner, disc_r_outer, disc_peak_T, aspect;
    T h_step, escape_r, horizon_cut, time_sec;
    std::uint32_t width, height, metric_type, flags, max_steps, frame_index, supersample;
    T exposure, bloom_threshold, bloom_strength, disc_turbulence;
    Extras<T> ext;
};

template <class T> inline Uni<T> uni_from_raw(const UniformsRaw& u) {
    Uni<T> o;
    o.cam_pos = v3<T>(u.cam_pos.x, u.cam_pos.y, u.cam_pos.z);
    o.cam_fwd = v3<T>(u.cam_fwd.x, u.cam_fwd.y, u.cam_fwd.z);
    o.cam_right = v3<T>(u.cam_right.x, u.cam_right.y, u.cam_right.z);
    o.cam_up = v3<T>(u.cam_up.x, u.cam_up.y, u.cam_up.z);
    o.mass_M = u.mass_M; o.spin_a = u.spin_a; o.rs = u.rs; o.tan_half_fov = u.tan_half_fov;
    o.disc_r_inner = u.disc_r_inner; o.disc_r_outer = u.disc_r_outer;
    o.disc_peak_T = u.disc_peak_T; o.aspect = u.aspect;
    o.h_step = u.h_step; o.escape_r = u.escape_r; o.horizon_cut = u.horizon_cut; o.time_sec = u.time_sec;
    o.width = u.width; o.height = u.height; o.metric_type = u.metric_type; o.flags = u.flags;
    o.max_steps = u.max_steps; o.frame_index = u.frame_index; o.supersample = u.supersample;
    o.exposure = u.exposure; o.bloom_threshold = u.bloom_threshold;
    o.bloom_strength = u.bloom_strength; o.disc_turbulence = u.disc_turbulence;
    return o;
}

// ---------------------------------------------------------------------------
// Kerr closed forms (core/include/physics/kerr.hpp:36-108).
// ---------------------------------------------------------------------------
template <class T> inline T kerr_r_plus(T M, T a) {  // kerr.hpp:36-39
    const T d = M * M - a * a;
    return M + xsqrt(d > T(0.0) ? d : T(0.0));
}
template <class T> inline T kerr_r_minus(T M, T a) {  // kerr.hpp:44-47
    const T d = M * M - a * a;
    return M - xsqrt(d > T(0.0) ? d : T(0.0));
}
template <class T> inline T kerr_photon_pro(T M, T a) {  // kerr.hpp:64-67
    return T(2.0) * M * (T(1.0) + xcos((T(2.0) / T(3.0)) * xacos(-a / M)));
}
template <class T> inline T kerr_photon_ret(T M, T a) {  // kerr.hpp:72-75
    return T(2.0) * M * (T(1.0) + xcos((T(2.0) / T(3.0)) * xacos(a / M)));
}
template <class T> inline T kerr_isco(T M, T a, bool prograde) {  // kerr.hpp:83-96
    const T x = a / M;
    const T x2 = x * x;
    const T z1 = T(1.0) + xcbrt(T(1.0) - x2) * (xcbrt(T(1.0) + x) + xcbrt(T(1.0) - x));
    const T z2 = xsqrt(T(3.0) * x2 + z1 * z1);
    const T rad = (T(3.0) - z1) * (T(3.0) + z1 + T(2.0) * z2);
    const T root = xsqrt(rad > T(0.0) ? rad : T(0.0));
    return M * (T(3.0) + z2 + (prograde ? -root : root));
}

// ---------------------------------------------------------------------------
// Shading helpers (render/cuda/kernels/geodesic_kernel.cu:77-171, 413-451).
// ---------------------------------------------------------------------------
template <class T> inline V3<T> blackbody_srgb(T Tk) {  // geodesic_kernel.cu:77-97
    Tk = vclamp(Tk, T(1000.0), T(40000.0)) * T(0.01);
    T r, g, b;
    if (Tk <= T(66.0)) {
        r = T(1.0);
        g = vclamp(T(0.39008157876901960784) * xlog(Tk) - T(0.63184144378862745098), T(0.0), T(1.0));
    } else {
        const T t1 = Tk - T(60.0);
        r = vclamp(T(1.29293618606274509804) * xpow(t1, T(-0.1332047592)), T(0.0), T(1.0));
        g = vclamp(T(1.12989086089529411765) * xpow(t1, T(-0.0755148492)), T(0.0), T(1.0));
    }
    if (Tk >= T(66.0)) {
        b = T(1.0);
    } else if (Tk <= T(19.0)) {
        b = T(0.0);
    } else {
        const T t1 = Tk - T(10.0);
        b = vclamp(T(0.54320678911019607843) * xlog(t1) - T(1.19625408914), T(0.0), T(1.0));
    }
    return v3(r, g, b);
}

template <class T> inline T disc_temperature(T r_cross, T r_inner, T peak_T) {  // geodesic_kernel.cu:99-107
    if (r_cross < r_inner) return T(0.0);
    const T ratio = r_inner / r_cross;
    const T q = xpow(ratio, T(0.75));
    const T edge = T(1.0) - xsqrt(ratio);
    const T raw = q * xfmax(edge, T(0.0));
    return peak_T * raw / T(0.186);
}

// geodesic_kernel.cu:109-121, plus the seeded phase offsets (0 = reference).
template <class T>
inline T disc_band_mask(T r_cross, T phi_cross, T r_inner, T turbulence, T time_sec, T ph1, T ph2) {
    const T rho = xlog(xfmax(r_cross / r_inner, T(0.1)));
    const T shear = T(4.0) / xfmax(xpow(r_cross / r_inner, T(1.5)), T(0.25));
    const T omega = xpow(xfmax(r_cross / r_inner, T(0.1)), T(-1.5));
    const T phi_anim = phi_cross - T(0.8) * time_sec * omega;
    const T band1 = T(0.5) + T(0.5) * xcos(T(10.0) * rho + T(3.0) * phi_anim + shear + ph1);
    const T band2 = T(0.5) + T(0.5) * xcos(T(22.0) * rho - T(1.7) * phi_anim + ph2);
    const T combined = vmix(T(0.8), band1 * band1 * band2, turbulence);
    return vclamp(T(0.65) + T(0.55) * combined, T(0.0), T(1.3));
}

inline std::uint32_t xorshift32(std::uint32_t x) {  // geodesic_kernel.cu:123-128
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}
inline std::uint32_t mix32(std::uint32_t a, std::uint32_t b) {  // geodesic_kernel.cu:129-131
    return xorshift32(a * 0x9E3779B9u + b);
}

template <class T> inline V3<T> starfield_color(V3<T> dir) {  // geodesic_kernel.cu:135-171
    const T len = vlength(dir);
    if (len < T(1e-6)) return v3(T(0.0), T(0.0), T(0.005));
    const T il = T(1.0) / len;
    const V3<T> n = dir * il;
    const T pi = T(3.14159265358979323846);
    const T theta = xacos(vclamp(n.z, T(-1.0), T(1.0)));
    const T phi = xatan2(n.y, n.x) + pi;
    const int NX = 768, NY = 384;
    int ix = int(val(phi / (T(2.0) * pi) * T(double(NX)))) % NX;
    int iy = int(val(theta / pi * T(double(NY)))) % NY;
    if (ix < 0) ix += NX;
    if (iy < 0) iy += NY;
    const std::uint32_t seed = mix32(std::uint32_t(ix), std::uint32_t(iy));
    T r = T(0.0015), g = T(0.002), b = T(0.005);
    if ((seed & 0xFFu) < 4u) {
        const T bright_raw = T(double((seed >> 8) & 0xFFu)) / T(255.0);
        const T brightness = xpow(bright_raw, T(4.0));
        const T hue = T(double((seed >> 16) & 0xFFu)) / T(255.0);
        const T sr = T(1.0) - T(0.5) * hue;
        const T sg = T(0.8) + T(0.2) * hue * (T(1.0) - hue) * T(4.0);
        const T sb = T(0.4) + T(0.6) * hue;
        r = xfmin(T(1.0), r + brightness * sr);
        g = xfmin(T(1.0), g + brightness * sg);
        b = xfmin(T(1.0), b + brightness * sb);
    }
    return v3(r, g, b) * T(0.6);
}

template <class T> inline V3<T> aces(V3<T> x) {  // geodesic_kernel.cu:413-422
    auto f = [](T c) {
        return vclamp((c * (T(2.51) * c + T(0.03))) / (c * (T(2.43) * c + T(0.59)) + T(0.14)), T(0.0), T(1.0));
    };
    return v3(f(x.x), f(x.y), f(x.z));
}
template <class T> inline T linear_to_srgb(T c) {  // geodesic_kernel.cu:424-427
    c = vclamp(c, T(0.0), T(1.0));
    return c <= T(0.0031308) ? T(12.92) * c : T(1.055) * xpow(c, T(1.0) / T(2.4)) - T(0.055);
}
template <class T> inline std::uint8_t to_u8(T c) {  // geodesic_kernel.cu:429-431
    return std::uint8_t(val(vclamp(linear_to_srgb(c) * T(255.0) + T(0.5), T(0.0), T(255.0))));
}
inline float halton(std::uint32_t i, std::uint32_t base) {  // geodesic_kernel.cu:442-451 (float)
    float f = 1.0f, r = 0.0f;
    while (i > 0u) {
        f /= float(base);
        r += f * float(i % base);
        i /= base;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Camera ray (geodesic_kernel.cu:177-223).  The camera-position part depends
// only on the scene; it is split out so callers can evaluate it once.
// ---------------------------------------------------------------------------
template <class T> struct CamFrame {
    T cam_r, cam_theta, cam_phi, sin_th, cos_th;
    V3<T> rhat, that, phat;
};
template <class T> inline CamFrame<T> camera_frame(const Uni<T>& u) {
    CamFrame<T> f;
    const V3<T> p = u.cam_pos;
    f.cam_r = vlength(p);
    f.cam_theta = xacos(vclamp(p.z / xfmax(f.cam_r, T(1e-6)), T(-1.0), T(1.0)));
    f.cam_phi = xatan2(p.y, p.x);
    f.sin_th = xsin(f.cam_theta);
    f.cos_th = xcos(f.cam_theta);
    const T cp = xcos(f.cam_phi), sp = xsin(f.cam_phi);
    f.rhat = v3(f.sin_th * cp, f.sin_th * sp, f.cos_th);
    f.that = v3(f.cos_th * cp, f.cos_th * sp, -f.sin_th);
    f.phat = v3(-sp, cp, T(0.0));
    return f;
}
template <class T> struct CameraRay {
    V3<T> dir;
    T ur, utheta, uphi;
};
template <class T>
inline CameraRay<T> camera_ray(const Uni<T>& u, const CamFrame<T>& f, std::uint32_t gx, std::uint32_t gy, T ox, T oy) {
    const T fx = T(double(gx)) + ox;
    const T fy = T(double(gy)) + oy;
    const T u_ndc = (T(2.0) * fx / T(double(u.width)) - T(1.0)) * u.aspect * u.tan_half_fov;
    const T v_ndc = -(T(2.0) * fy / T(double(u.height)) - T(1.0)) * u.tan_half_fov;
    CameraRay<T> r;
    r.dir = vnormalize(u.cam_fwd + (u.cam_right * u_ndc + u.cam_up * v_ndc));
    r.ur = vdot(r.dir, f.rhat);
    r.utheta = vdot(r.dir, f.that) / f.cam_r;
    r.uphi = vdot(r.dir, f.phat) / (f.cam_r * xfmax(f.sin_th, T(1e-6)));
    return r;
}

// ---------------------------------------------------------------------------
// Schwarzschild 8-component geodesic (shared_shader/geodesic_math.h:25-134).
// ---------------------------------------------------------------------------
template <class T> struct SState {
    T t, r, theta, phi, ut, ur, utheta, uphi;
};
template <class T> inline SState<T> s_axpy(const SState<T>& y, const SState<T>& k, T h) {  // y + k*h
    return {y.t + k.t * h, y.r + k.r * h, y.theta + k.theta * h, y.phi + k.phi * h,
            y.ut + k.ut * h, y.ur + k.ur * h, y.utheta + k.utheta * h, y.uphi + k.uphi * h};
}
template <class T> inline SState<T> schw_rhs(const SState<T>& s, T rs) {  // geodesic_math.h:101-123
    const T f = T(1.0) - rs / s.r;
    const T r2 = s.r * s.r;
    const T sn = xsin(s.theta);
    const T cs = xcos(s.theta);
    SState<T> d;
    d.t = s.ut;
    d.r = s.ur;
    d.theta = s.utheta;
    d.phi = s.uphi;
    d.ut = -(rs / (r2 * f)) * s.ut * s.ur;
    d.ur = -(rs * f / (T(2.0) * r2)) * s.ut * s.ut + (rs / (T(2.0) * r2 * f)) * s.ur * s.ur
           + s.r * f * (s.utheta * s.utheta + sn * sn * s.uphi * s.uphi);
    d.utheta = T(-2.0) * s.ur * s.utheta / s.r + sn * cs * s.uphi * s.uphi;
    d.uphi = T(-2.0) * s.ur * s.uphi / s.r - T(2.0) * (cs / sn) * s.utheta * s.uphi;
    return d;
}
template <class T> inline SState<T> schw_rk4(const SState<T>& y, T h, T rs) {  // geodesic_math.h:126-134
    const SState<T> k1 = schw_rhs(y, rs);
    const SState<T> k2 = schw_rhs(s_axpy(y, k1, T(0.5) * h), rs);
    const SState<T> k3 = schw_rhs(s_axpy(y, k2, T(0.5) * h), rs);
    const SState<T> k4 = schw_rhs(s_axpy(y, k3, h), rs);
    SState<T> sum;
    sum.t = (k1.t + k2.t * T(2.0)) + (k3.t * T(2.0) + k4.t);
    sum.r = (k1.r + k2.r * T(2.0)) + (k3.r * T(2.0) + k4.r);
    sum.theta = (k1.theta + k2.theta * T(2.0)) + (k3.theta * T(2.0) + k4.theta);
    sum.phi = (k1.phi + k2.phi * T(2.0)) + (k3.phi * T(2.0) + k4.phi);
    sum.ut = (k1.ut + k2.ut * T(2.0)) + (k3.ut * T(2.0) + k4.ut);
    sum.ur = (k1.ur + k2.ur * T(2.0)) + (k3.ur * T(2.0) + k4.ur);
    sum.utheta = (k1.utheta + k2.utheta * T(2.0)) + (k3.utheta * T(2.0) + k4.utheta);
    sum.uphi = (k1.uphi + k2.uphi * T(2.0)) + (k3.uphi * T(2.0) + k4.uphi);
    return s_axpy(y, sum, h / T(6.0));
}

// ---------------------------------------------------------------------------
// Kerr Hamiltonian form (shared_shader/kerr_hamilton.h:37-178, kerr_math.h:46-116).
// ---------------------------------------------------------------------------
template <class T> struct KConst {
    T E, L_z, Q, a, M;
};
template <class T> struct KState {
    T t, r, theta, phi, p_r, p_theta;
};
template <class T> inline KState<T> k_axpy(const KState<T>& y, const KState<T>& k, T h) {
    return {y.t + k.t * h, y.r + k.r * h, y.theta + k.theta * h, y.phi + k.phi * h,
            y.p_r + k.p_r * h, y.p_theta + k.p_theta * h};
}
template <class T> inline T kerr_two_H(T r, T theta, T p_r, T p_theta, const KConst<T>& c) {  // kerr_hamilton.h:67-94
    const T st = xsin(theta);
    const T s2 = st * st;
    const T safe_s2 = (s2 < T(1e-10)) ? T(1e-10) : s2;
    const T ct = xcos(theta);
    const T c2 = ct * ct;
    const T r2 = r * r;
    const T a = c.a, M = c.M;
    const T a2 = a * a;
    const T Sigma = r2 + a2 * c2;
    const T Delta = r2 - T(2.0) * M * r + a2;
    const T A = (r2 + a2) * (r2 + a2) - a2 * Delta * s2;
    const T SD = Sigma * Delta;
    const T t_tt = -A * c.E * c.E / SD;
    const T t_tp = T(4.0) * M * r * a * c.E * c.L_z / SD;
    const T t_pp = (Delta - a2 * safe_s2) * c.L_z * c.L_z / (SD * safe_s2);
    const T t_rr = Delta * p_r * p_r / Sigma;
    const T t_th = p_theta * p_theta / Sigma;
    return t_tt + t_tp + t_pp + t_rr + t_th;
}
template <class T> inline KState<T> kerr_rhs(const KState<T>& s, const KConst<T>& c) {  // kerr_hamilton.h:97-155
    const T r = s.r;
    const T st = xsin(s.theta);
    const T s2 = st * st;
    const T safe_s2 = (s2 < T(1e-10)) ? T(1e-10) : s2;
    const T ct = xcos(s.theta);
    const T c2 = ct * ct;
    const T r2 = r * r;
    const T a = c.a, M = c.M;
    const T a2 = a * a;
    const T Sigma = r2 + a2 * c2;
    const T Delta = r2 - T(2.0) * M * r + a2;
    const T A = (r2 + a2) * (r2 + a2) - a2 * Delta * s2;
    const T SD = Sigma * Delta;
    KState<T> d;
    d.t = (A * c.E - T(2.0) * M * r * a * c.L_z) / SD;
    d.phi = (T(2.0) * M * r * a * c.E * s2 + (Delta - a2 * s2) * c.L_z) / (SD * safe_s2);
    d.r = Delta * s.p_r / Sigma;
    d.theta = s.p_theta / Sigma;
    const T E = c.E, L = c.L_z, L2 = L * L;
    const T B = -A * E * E + T(4.0) * M * r * a * E * L - a2 * L2;
    const T Sig_r = T(2.0) * r;
    const T Del_r = T(2.0) * r - T(2.0) * M;
    const T A_r = T(4.0) * r * (r2 + a2) - a2 * Del_r * s2;
    const T B_r = -A_r * E * E + T(4.0) * M * a * E * L;
    const T sin2t = T(2.0) * st * ct;
    const T Sig_t = -a2 * sin2t;
    const T B_t = a2 * Delta * sin2t * E * E;
    const T Sigma2 = Sigma * Sigma;
    const T Dp2 = Delta * s.p_r * s.p_r + s.p_theta * s.p_theta;
    const T dH_dr = B_r / SD - B * (Sig_r * Delta + Sigma * Del_r) / (SD * SD) - L2 * Sig_r / (Sigma2 * safe_s2)
                    + Del_r * s.p_r * s.p_r / Sigma - Dp2 * Sig_r / Sigma2;
    const T dH_dth = B_t / SD - B * Sig_t / (Sigma2 * Delta)
                     - L2 * (Sig_t * safe_s2 + Sigma * sin2t) / (Sigma2 * safe_s2 * safe_s2) - Dp2 * Sig_t / Sigma2;
    d.p_r = T(-0.5) * dH_dr;
    d.p_theta = T(-0.5) * dH_dth;
    return d;
}
template <class T> inline T kerr_pole_scale(T theta) {  // kerr_hamilton.h:166-169
    T st_abs = xsin(theta);
    st_abs = (st_abs < T(0.0)) ? -st_abs : st_abs;
    T ps = st_abs * T(5.0);
    ps = (ps > T(1.0)) ? T(1.0) : ((ps < T(0.05)) ? T(0.05) : ps);
    return ps;
}
template <class T> inline KState<T> kerr_rk4_raw(const KState<T>& y, T h, const KConst<T>& c) {  // kerr_hamilton.h:171-177
    const KState<T> k1 = kerr_rhs(y, c);
    const KState<T> k2 = kerr_rhs(k_axpy(y, k1, T(0.5) * h), c);
    const KState<T> k3 = kerr_rhs(k_axpy(y, k2, T(0.5) * h), c);
    const KState<T> k4 = kerr_rhs(k_axpy(y, k3, h), c);
    KState<T> sum;
    sum.t = (k1.t + k2.t * T(2.0)) + (k3.t * T(2.0) + k4.t);
    sum.r = (k1.r + k2.r * T(2.0)) + (k3.r * T(2.0) + k4.r);
    sum.theta = (k1.theta + k2.theta * T(2.0)) + (k3.theta * T(2.0) + k4.theta);
    sum.phi = (k1.phi + k2.phi * T(2.0)) + (k3.phi * T(2.0) + k4.phi);
    sum.p_r = (k1.p_r + k2.p_r * T(2.0)) + (k3.p_r * T(2.0) + k4.p_r);
    sum.p_theta = (k1.p_theta + k2.p_theta * T(2.0)) + (k3.p_theta * T(2.0) + k4.p_theta);
    return k_axpy(y, sum, h / T(6.0));
}
template <class T> inline KState<T> kerr_rk4(const KState<T>& y, T h, const KConst<T>& c) {  // kerr_hamilton.h:165-178
    return kerr_rk4_raw(y, h * kerr_pole_scale(y.theta), c);
}
// Carter constant from (theta, p_theta): Q = p_theta^2 + cos^2(L^2/sin^2 - a^2E^2)
// (kerr_math.h:107-116 Theta-potential rearranged; geodesic_kernel.cu:335-336).
template <class T> inline T kerr_carter_Q(T theta, T p_theta, const KConst<T>& c) {
    const T s = xsin(theta), cth = xcos(theta);
    const T s2 = s * s;
    const T safe = (s2 < T(1e-12)) ? T(1e-12) : s2;
    return p_theta * p_theta + cth * cth * (c.L_z * c.L_z / safe - c.a * c.a * c.E * c.E);
}

// ---------------------------------------------------------------------------
// Equatorial crossing (shared_shader/disc_intersection.h:23-55).
// ---------------------------------------------------------------------------
template <class T> struct Crossing {
    bool crossed;
    T frac;
};
template <class T> inline Crossing<T> detect_crossing(T th_prev, T th_curr) {
    const T mid = T(1.5707963267948966192313216916397514);  // pi/2 in T (float literal in FP32)
    const T dp = th_prev - mid;
    const T dc = th_curr - mid;
    Crossing<T> x;
    x.crossed = (dp * dc) < T(0.0);
    const T denom = th_curr - th_prev;
    x.frac = x.crossed ? (mid - th_prev) / denom : T(0.0);
    return x;
}
template <class T> inline T lerp_scalar(T a, T b, T f) { return a + f * (b - a); }
template <class T> inline bool in_annulus(T r, T r_in, T r_out) { return (r >= r_in) && (r <= r_out); }

// ---------------------------------------------------------------------------
// Per-ray trace.  Termination classes.
// ---------------------------------------------------------------------------
enum RayClass : std::uint8_t { CLS_HORIZON = 0, CLS_ESCAPE = 1, CLS_DISC = 2, CLS_BUDGET = 3, CLS_INITFAIL = 4 };

template <class T> struct RayOut {
    V3<T> rgb{T(0.0), T(0.0), T(0.0)};
    int cls = CLS_BUDGET;
    std::uint32_t steps = 0;
    T r_cross = T(0.0);
    // diagnostics (value-level): invariants at start and at termination
    double inv0[3] = {0, 0, 0};  // Kerr: 2H, Q, - ; Schw: E, L, null residual
    double inv1[3] = {0, 0, 0};
    double E = 0, Lz = 0;        // conserved parameters at init
    double r_min = 1e30;         // smallest radius reached
    double th_min = 1e30;        // smallest |theta - pi/2| over the states after the first
    int n_soft = 0;              // crossings with nonzero opacity (soft-edge model)
    double esc_theta = 0, esc_phi = 0;  // direction the escape colour is looked up at
    double th_max = 0;           // largest |theta - pi/2| over the states after the first (pole proximity)
};

template <class T> inline T soft_alpha(T r, T r_in, T r_out, T w) {
    // Opacity of the differentiable disc model: 1 inside [r_in+w, r_out-w], ramping
    // linearly to 0 at r_in and r_out (zero outside the annulus, so the emission's
    // jump at r_in — the 1000 K floor of the blackbody — is multiplied by 0).
    // w = 0 is never evaluated (hard edges).
    const T lo = (r - r_in) / w;
    const T hi = (r_out - r) / w;
    return vclamp(xfmin(lo, hi), T(0.0), T(1.0));
}

// Frequency-shift factors at a disc crossing.  g_grav and g_dop follow the
// kernel's Schwarzschild shading (geodesic_kernel.cu:266-279): the host twins
// are schwarzschild_grav_redshift, schwarzschild_keplerian_omega and
// schwarzschild_doppler_factor (core/include/physics/redshift.hpp:44-97),
// here with the kernel's clamps.
template <class T> struct Shift {
    T g_grav, g_dop, omega;
};
template <class T> inline Shift<T> schw_shift(T rs, T M, T r_cross, T uphi_cross, T f_cam, unsigned flags) {
    Shift<T> s{T(1.0), T(1.0), T(0.0)};
    const T f_emit = T(1.0) - rs / r_cross;
    if (flags & FLAG_REDSHIFT) s.g_grav = xsqrt(xfmax(f_emit / f_cam, T(0.0)));
    if (flags & FLAG_DOPPLER) {
        s.omega = xsqrt(M / (r_cross * r_cross * r_cross));
        const T v = r_cross * s.omega / xsqrt(xfmax(f_emit, T(1e-6)));
        const T cos_psi = vclamp(-uphi_cross * r_cross, T(-1.0), T(1.0));
        s.g_dop = xsqrt(xfmax(T(1.0) - v * v, T(0.0))) / xfmax(T(1.0) - v * cos_psi, T(1e-4));
    }
    return s;
}
// Kerr path (geodesic_kernel.cu:368-386): Schwarzschild lapse ratio for g_grav,
// prograde Keplerian omega_K = omega_0 / (1 + a omega_0), cos psi = L_z / r.
template <class T> inline Shift<T> kerr_shift(T M, T a, T r_cross, T cam_r, T L_z, unsigned flags) {
    Shift<T> s{T(1.0), T(1.0), T(0.0)};
    const T f_emit = T(1.0) - T(2.0) * M / r_cross;
    const T f_obs = T(1.0) - T(2.0) * M / cam_r;
    if (flags & FLAG_REDSHIFT) s.g_grav = xsqrt(xfmax(f_emit, T(1e-6)) / xfmax(f_obs, T(1e-6)));
    if (flags & FLAG_DOPPLER) {
        const T om0 = xsqrt(M / (r_cross * r_cross * r_cross));
        s.omega = om0 / (T(1.0) + a * om0);
        const T v = r_cross * s.omega / xsqrt(xfmax(f_emit, T(1e-6)));
        const T cos_psi = vclamp(L_z / r_cross, T(-1.0), T(1.0));
        const T v2 = v * v;
        s.g_dop = xsqrt(xfmax(T(1.0) - v2, T(0.0))) / xfmax(T(1.0) - v * cos_psi, T(1e-4));
    }
    return s;
}
// Flux clamp of the disc emission (geodesic_kernel.cu:281,388).
template <class T> inline T disc_flux(T r_in, T r_cross) {
    return vclamp(xpow(r_in / r_cross, T(3.0)), T(0.05), T(1.5));
}
// blackbody(g T) x flux x bands x g^4 (geodesic_kernel.cu:280-286,387-393; g^4 is
// intensity_scaling, redshift.hpp:106-109).
template <class T> inline V3<T> disc_emit(const Uni<T>& u, const Shift<T>& sh, T r_cross, T phi_cross, T Tk) {
    const T g = sh.g_grav * sh.g_dop;
    const V3<T> col = blackbody_srgb(g * Tk);
    const T flux = disc_flux(u.disc_r_inner, r_cross);
    const T bands = disc_band_mask(r_cross, phi_cross, u.disc_r_inner, u.disc_turbulence, u.time_sec,
                                   u.ext.band_phase1, u.ext.band_phase2);
    const T g2 = g * g;
    return col * (flux * bands * g2 * g2);
}

// Disc emission at a crossing for the Schwarzschild path (geodesic_kernel.cu:264-289).
template <class T> inline V3<T> schw_disc_emit(const Uni<T>& u, T r_cross, T uphi_cross, T phi_cross, T f_cam, T Tk) {
    return disc_emit(u, schw_shift(u.rs, u.mass_M, r_cross, uphi_cross, f_cam, u.flags), r_cross, phi_cross, Tk);
}
// Disc emission at a crossing for the Kerr path (geodesic_kernel.cu:366-394).
template <class T> inline V3<T> kerr_disc_emit(const Uni<T>& u, T r_cross, T phi_cross, T cam_r, T L_z, T Tk) {
    return disc_emit(u, kerr_shift(u.mass_M, u.spin_a, r_cross, cam_r, L_z, u.flags), r_cross, phi_cross, Tk);
}

template <class T> inline V3<T> escape_color(const Uni<T>& u, T theta, T phi) {  // geodesic_kernel.cu:251-259
    if (u.flags & FLAG_STARFIELD) {
        const T sx = xsin(theta) * xcos(phi);
        const T sy = xsin(theta) * xsin(phi);
        const T sz = xcos(theta);
        return starfield_color(v3(sx, sy, sz));
    }
    return v3(T(0.0), T(0.0), T(0.0));
}

// Disc hit bookkeeping shared by both metrics, with the optional soft edge.
// Returns true when the ray terminates at this crossing.
template <class T, class EmitFn>
inline bool disc_hit(const Uni<T>& u, T r_cross, V3<T>& acc, T& trans, RayOut<T>& o, EmitFn emit) {
    const T w = u.ext.soft_edge;
    const T peak = u.disc_peak_T * u.ext.peak_T_scale;
    if (val(w) <= 0.0) {
        if (!in_annulus(r_cross, u.disc_r_inner, u.disc_r_outer)) return false;
        const T Tk = disc_temperature(r_cross, u.disc_r_inner, peak);
        if (Tk > T(0.0)) acc = acc + emit(Tk) * trans;
        o.cls = CLS_DISC;
        o.r_cross = r_cross;
        return true;
    }
    const T alpha = soft_alpha(r_cross, u.disc_r_inner, u.disc_r_outer, w);
    if (!(alpha > T(0.0))) return false;
    const T Tk = disc_temperature(r_cross, u.disc_r_inner, peak);
    if (Tk > T(0.0)) acc = acc + emit(Tk) * (trans * alpha);
    o.r_cross = r_cross;
    o.n_soft += 1;
    // opaque crossing, or the 8th crossing of the soft model: the ray ends here
    if (!(val(alpha) < 1.0) || o.n_soft == 8) {
        o.cls = CLS_DISC;
        trans = T(0.0);
        return true;
    }
    trans = trans * (T(1.0) - alpha);
    return false;
}

// One RK4 step of either metric with the optional error control (step
// doubling).  Returns the step actually taken in h_used; h_next is the
// controller's proposal.  With adapt_tol == 0 this is the reference's fixed step.
template <class T>
inline KState<T> kerr_step(const KState<T>& y, const KConst<T>& c, const Uni<T>& u, T& h_cur, T* h_used = nullptr) {
    if (!(val(u.ext.adapt_tol) > 0.0)) return kerr_rk4(y, u.h_step, c);
    // Step doubling on the damped step: one h step vs two h/2 steps.
    const T tol = u.ext.adapt_tol;
    for (int attempt = 0; attempt < 12; ++attempt) {
        const T h = h_cur * kerr_pole_scale(y.theta);
        const KState<T> big = kerr_rk4_raw(y, h, c);
        const KState<T> half = kerr_rk4_raw(kerr_rk4_raw(y, T(0.5) * h, c), T(0.5) * h, c);
        // error norm: max over (r, theta, phi) scaled by (1 + |r|) for r.
        const T er = (half.r - big.r) / (T(1.0) + (half.r < T(0.0) ? -half.r : half.r));
        const T et = half.theta - big.theta;
        const T ep = half.phi - big.phi;
        T err = xfmax(xfmax(er < T(0.0) ? -er : er, et < T(0.0) ? -et : et), ep < T(0.0) ? -ep : ep) / T(15.0);
        if (!(err <= tol) && attempt < 11 && h_cur > u.h_step * T(1e-3)) {
            h_cur = h_cur * T(0.5);
            continue;
        }
        T grow = (err > T(0.0)) ? T(0.9) * xpow(tol / err, T(0.2)) : T(4.0);
        grow = vclamp(grow, T(0.5), T(4.0));
        h_cur = xfmin(h_cur * grow, u.h_step * T(16.0));
        if (h_used) *h_used = h;
        return half;
    }
    return kerr_rk4(y, h_cur, c);
}

// Equator crossing of an error-controlled step (no counterpart in the
// reference, whose fixed steps interpolate linearly): cubic Hermite dense
// output on [y0, y1] from the states and their RHS derivatives (step h), three
// Newton iterations on theta(s) = pi/2 from the linear estimate, then r and phi
// at that s.  Accurate to O(h^4), where the linear form is O(h^2): error-
// controlled steps are long (up to 16 h_step) far from the hole.
template <class T> struct DenseCross {
    T r, phi;
};
template <class T> inline T hermite(T s, T x0, T d0, T x1, T d1) {
    const T s2 = s * s, s3 = s2 * s;
    return (T(2.0) * s3 - T(3.0) * s2 + T(1.0)) * x0 + (s3 - T(2.0) * s2 + s) * d0 + (T(3.0) * s2 - T(2.0) * s3) * x1 +
           (s3 - s2) * d1;
}
template <class T> inline T hermite_ds(T s, T x0, T d0, T x1, T d1) {
    const T s2 = s * s;
    return (T(6.0) * s2 - T(6.0) * s) * (x0 - x1) + (T(3.0) * s2 - T(4.0) * s + T(1.0)) * d0 + (T(3.0) * s2 - T(2.0) * s) * d1;
}
// Three Newton iterations on hermite(s) = target from s, clamped to [0, 1].
template <class T> inline T dense_root(T x0, T d0, T x1, T d1, T target, T s) {
    for (int it = 0; it < 3; ++it) {
        const T g = hermite(s, x0, d0, x1, d1) - target;
        const T dg = hermite_ds(s, x0, d0, x1, d1);
        if (val(dg) != 0.0) s = s - g / dg;
        s = vclamp(s, T(0.0), T(1.0));
    }
    return s;
}
template <class T>
inline DenseCross<T> kerr_dense_crossing(const KState<T>& y0, const KState<T>& y1, T h, T frac, const KConst<T>& c) {
    const KState<T> f0 = kerr_rhs(y0, c), f1 = kerr_rhs(y1, c);
    const T mid = T(1.5707963267948966192313216916397514);
    const T s = dense_root(y0.theta, h * f0.theta, y1.theta, h * f1.theta, mid, frac);
    return {hermite(s, y0.r, h * f0.r, y1.r, h * f1.r), hermite(s, y0.phi, h * f0.phi, y1.phi, h * f1.phi)};
}
// Null-constraint re-projection (an option with no counterpart in the reference,
// whose constraint kerr_ham_two_H, kerr_hamilton.h:67-94, only its tests
// evaluate): one Newton step on 2H(r, theta, p_r, p_theta) = 0 along the gradient
// of 2H in the momenta, g = (2 Delta p_r, 2 p_theta) / Sigma, i.e.
// p -= (2H / |g|^2) g.  2H is quadratic in (p_r, p_theta), so what remains is
// O((2H)^2); r, theta, phi, E and L_z are untouched.
template <class T> inline KState<T> kerr_reproject(KState<T> s, const KConst<T>& c) {
    const T ct = xcos(s.theta);
    const T r2 = s.r * s.r, a2 = c.a * c.a;
    const T Sigma = r2 + a2 * ct * ct;
    const T Delta = r2 - T(2.0) * c.M * s.r + a2;
    const T h2 = kerr_two_H(s.r, s.theta, s.p_r, s.p_theta, c);
    const T gr = T(2.0) * Delta * s.p_r / Sigma, gt = T(2.0) * s.p_theta / Sigma;
    const T n = gr * gr + gt * gt;
    if (val(n) > 0.0) {
        const T lam = h2 / n;
        s.p_r = s.p_r - lam * gr;
        s.p_theta = s.p_theta - lam * gt;
    }
    return s;
}
// The precision a crossing refinement runs in: double for a float march.
template <class T> struct Wide { using type = T; };
template <> struct Wide<float> { using type = double; };
// Equator crossing of a fixed step with the root refined in FP64 (an option; the
// reference interpolates linearly in theta, disc_intersection.h:30-44): the
// cubic Hermite dense output of the step the march took (h = h_step *
// pole_scale(theta_prev)) from its end states and their RHS derivatives in the
// march's precision T, then, widened to FP64, three Newton iterations on
// theta(s) = pi/2 from the linear fraction and r, phi at the root.
template <class T>
inline DenseCross<T> kerr_refined_crossing(const KState<T>& y0, const KState<T>& y1, T h, const KConst<T>& c) {
    using W = typename Wide<T>::type;
    const KState<T> f0 = kerr_rhs(y0, c), f1 = kerr_rhs(y1, c);
    const W mid = W(1.5707963267948966192313216916397514);
    const W th0 = W(y0.theta), th1 = W(y1.theta);
    const W s = dense_root(th0, W(h * f0.theta), th1, W(h * f1.theta), mid, (mid - th0) / (th1 - th0));
    return {T(hermite(s, W(y0.r), W(h * f0.r), W(y1.r), W(h * f1.r))),
            T(hermite(s, W(y0.phi), W(h * f0.phi), W(y1.phi), W(h * f1.phi)))};
}

// Escape of an error-controlled step: the direction (theta, phi) where the dense
// output reaches r = escape_r (the fixed-step algorithm takes the first state
// beyond it, which moves by a whole step when the step sequence changes).
template <class T> struct DenseEscape {
    T theta, phi;
};
template <class T>
inline DenseEscape<T> kerr_dense_escape(const KState<T>& y0, const KState<T>& y1, T h, T r_esc, const KConst<T>& c) {
    const KState<T> f0 = kerr_rhs(y0, c), f1 = kerr_rhs(y1, c);
    const T s = dense_root(y0.r, h * f0.r, y1.r, h * f1.r, r_esc, (r_esc - y0.r) / (y1.r - y0.r));
    return {hermite(s, y0.theta, h * f0.theta, y1.theta, h * f1.theta),
            hermite(s, y0.phi, h * f0.phi, y1.phi, h * f1.phi)};
}

// trace_kerr (geodesic_kernel.cu:303-403) with invariants recorded.
template <class T>
inline RayOut<T> trace_kerr(const Uni<T>& u, const CamFrame<T>& f, const CameraRay<T>& cam) {
    RayOut<T> o;
    const T M = u.mass_M, a = u.spin_a;
    const T sin_th = f.sin_th;
    const T cos_th = xcos(f.cam_theta);
    const T cam_r = f.cam_r;
    const T Sigma_c = cam_r * cam_r + a * a * cos_th * cos_th;
    const T Delta_c = cam_r * cam_r - T(2.0) * M * cam_r + a * a;
    const T s2_c = sin_th * sin_th;
    const T g_tt = -(T(1.0) - T(2.0) * M * cam_r / Sigma_c);
    const T g_tp = T(-2.0) * M * cam_r * a * s2_c / Sigma_c;
    const T g_rr = Sigma_c / Delta_c;
    const T g_thth = Sigma_c;
    const T g_pp = s2_c * ((cam_r * cam_r + a * a) * (cam_r * cam_r + a * a) - a * a * Delta_c * s2_c) / Sigma_c;
    const T A_q = g_tt;
    const T B_q = T(2.0) * g_tp * cam.uphi;
    const T C_q = g_rr * cam.ur * cam.ur + g_thth * cam.utheta * cam.utheta + g_pp * cam.uphi * cam.uphi;
    const T disc_q = B_q * B_q - T(4.0) * A_q * C_q;
    if (disc_q < T(0.0)) {
        o.cls = CLS_INITFAIL;
        return o;
    }
    const T ut = (-B_q - xsqrt(disc_q)) / (T(2.0) * A_q);
    KConst<T> c;
    c.a = a;
    c.M = M;
    c.E = -(g_tt * ut + g_tp * cam.uphi);
    c.L_z = g_tp * ut + g_pp * cam.uphi;
    const T p_theta_init = g_thth * cam.utheta;
    c.Q = p_theta_init * p_theta_init + cos_th * cos_th * (c.L_z * c.L_z / xfmax(s2_c, T(1e-6)) - a * a * c.E * c.E);
    KState<T> s;
    s.t = T(0.0);
    s.r = cam_r;
    s.theta = f.cam_theta;
    s.phi = f.cam_phi;
    s.p_r = g_rr * cam.ur;
    s.p_theta = g_thth * cam.utheta;
    o.E = val(c.E);
    o.Lz = val(c.L_z);
    o.inv0[0] = val(kerr_two_H(s.r, s.theta, s.p_r, s.p_theta, c));
    o.inv0[1] = val(kerr_carter_Q(s.theta, s.p_theta, c));
    T prev_theta = s.theta, prev_r = s.r, prev_phi = s.phi;
    KState<T> prev = s;  // error-controlled stepping: full state before the step, and its size
    T h_used = u.h_step;
    const bool adaptive = val(u.ext.adapt_tol) > 0.0;
    const T cut = T(1.02) * (M + xsqrt(xfmax(M * M - a * a, T(0.0))));
    V3<T> acc = v3(T(0.0), T(0.0), T(0.0));
    T trans = T(1.0);
    T h_cur = u.h_step;
    auto finish = [&](int cls) {
        o.cls = cls;
        o.rgb = acc;
        o.inv1[0] = val(kerr_two_H(s.r, s.theta, s.p_r, s.p_theta, c));
        o.inv1[1] = val(kerr_carter_Q(s.theta, s.p_theta, c));
        return o;
    };
    for (std::uint32_t step = 0; step < u.max_steps; ++step) {
        if (!xisfinite(s.r) || s.r < cut) return finish(CLS_HORIZON);
        if (s.r > u.escape_r) {
            T th_e = s.theta, ph_e = s.phi;
            if (adaptive && o.steps > 0) {
                const DenseEscape<T> de = kerr_dense_escape(prev, s, h_used, u.escape_r, c);
                th_e = de.theta;
                ph_e = de.phi;
            }
            o.esc_theta = val(th_e);
            o.esc_phi = val(ph_e);
            acc = acc + escape_color(u, th_e, ph_e) * trans;
            return finish(CLS_ESCAPE);
        }
        const Crossing<T> x = detect_crossing(prev_theta, s.theta);
        if ((u.flags & FLAG_DISC) && x.crossed) {
            T r_cross = lerp_scalar(prev_r, s.r, x.frac);
            // the reference shades at phi of the step after the crossing; the soft
            // (differentiable) model interpolates it like r; error-controlled steps
            // use the dense output for both
            T phi_cross = (val(u.ext.soft_edge) > 0.0) ? lerp_scalar(prev_phi, s.phi, x.frac) : s.phi;
            if (adaptive) {
                const DenseCross<T> dc = kerr_dense_crossing(prev, s, h_used, x.frac, c);
                r_cross = dc.r;
                phi_cross = dc.phi;
            } else if (u.ext.cross_refine && o.steps > 0) {
                const DenseCross<T> dc = kerr_refined_crossing(prev, s, u.h_step * kerr_pole_scale(prev.theta), c);
                r_cross = dc.r;
                phi_cross = dc.phi;
            }
            const bool stop = disc_hit(u, r_cross, acc, trans, o, [&](T Tk) {
                return kerr_disc_emit(u, r_cross, phi_cross, cam_r, c.L_z, Tk);
            });
            if (stop) return finish(CLS_DISC);
        }
        prev_theta = s.theta;
        prev_r = s.r;
        prev_phi = s.phi;
        prev = s;
        s = kerr_step(s, c, u, h_cur, &h_used);
        if (u.ext.reproject && !adaptive) s = kerr_reproject(s, c);
        o.steps = step + 1;
        o.r_min = std::min(o.r_min, val(s.r));
        o.th_min = std::min(o.th_min, std::fabs(val(s.theta) - 1.5707963267948966));
        o.th_max = std::max(o.th_max, std::fabs(val(s.theta) - 1.5707963267948966));
    }
    return finish(CLS_BUDGET);
}

template <class T> inline double schw_null(const SState<T>& s, T rs) {
    const double r = val(s.r), f = 1.0 - val(rs) / r, sn = std::sin(val(s.theta));
    const double ut = val(s.ut), ur = val(s.ur), uth = val(s.utheta), uph = val(s.uphi);
    return -f * ut * ut + ur * ur / f + r * r * uth * uth + r * r * sn * sn * uph * uph;
}

// trace_schwarzschild (geodesic_kernel.cu:229-297) with invariants recorded.
template <class T>
inline RayOut<T> trace_schw(const Uni<T>& u, const CamFrame<T>& f, const CameraRay<T>& cam) {
    RayOut<T> o;
    SState<T> s;
    s.t = T(0.0);
    s.r = f.cam_r;
    s.theta = f.cam_theta;
    s.phi = f.cam_phi;
    s.ur = cam.ur;
    s.utheta = cam.utheta;
    s.uphi = cam.uphi;
    const T f_cam = T(1.0) - u.rs / f.cam_r;
    const T kin = (cam.ur * cam.ur) / f_cam + cam.utheta * cam.utheta * f.cam_r * f.cam_r
                  + cam.uphi * cam.uphi * f.cam_r * f.cam_r * f.sin_th * f.sin_th;
    s.ut = xsqrt(xfmax(kin / f_cam, T(0.0)));
    auto invs = [&](double* out) {
        const double r = val(s.r), fr = 1.0 - val(u.rs) / r, sn = std::sin(val(s.theta));
        out[0] = fr * val(s.ut);
        out[1] = r * r * sn * sn * val(s.uphi);
        out[2] = schw_null(s, u.rs);
    };
    invs(o.inv0);
    o.E = o.inv0[0];
    o.Lz = o.inv0[1];
    T prev_theta = s.theta, prev_r = s.r, prev_uphi = s.uphi;
    V3<T> acc = v3(T(0.0), T(0.0), T(0.0));
    T trans = T(1.0);
    auto finish = [&](int cls) {
        o.cls = cls;
        o.rgb = acc;
        invs(o.inv1);
        return o;
    };
    for (std::uint32_t step = 0; step < u.max_steps; ++step) {
        if (!xisfinite(s.r) || s.r < u.horizon_cut) return finish(CLS_HORIZON);
        if (s.r > u.escape_r) {
            o.esc_theta = val(s.theta);
            o.esc_phi = val(s.phi);
            acc = acc + escape_color(u, s.theta, s.phi) * trans;
            return finish(CLS_ESCAPE);
        }
        const Crossing<T> x = detect_crossing(prev_theta, s.theta);
        if ((u.flags & FLAG_DISC) && x.crossed) {
            const T r_cross = lerp_scalar(prev_r, s.r, x.frac);
            const T uphi_cross = lerp_scalar(prev_uphi, s.uphi, x.frac);
            const T phi_cross = lerp_scalar(s.phi, s.phi, x.frac);  // reference quirk: lerp(phi, phi)
            const bool stop = disc_hit(u, r_cross, acc, trans, o, [&](T Tk) {
                return schw_disc_emit(u, r_cross, uphi_cross, phi_cross, f_cam, Tk);
            });
            if (stop) return finish(CLS_DISC);
        }
        prev_theta = s.theta;
        prev_r = s.r;
        prev_uphi = s.uphi;
        s = schw_rk4(s, u.h_step, u.rs);
        o.steps = step + 1;
        o.r_min = std::min(o.r_min, val(s.r));
        o.th_min = std::min(o.th_min, std::fabs(val(s.theta) - 1.5707963267948966));
        o.th_max = std::max(o.th_max, std::fabs(val(s.theta) - 1.5707963267948966));
    }
    return finish(CLS_BUDGET);
}

template <class T>
inline RayOut<T> trace_ray(const Uni<T>& u, const CamFrame<T>& f, std::uint32_t gx, std::uint32_t gy, T ox, T oy) {
    const CameraRay<T> cam = camera_ray(u, f, gx, gy, ox, oy);
    return (u.metric_type == METRIC_KERR) ? trace_kerr(u, f, cam) : trace_schw(u, f, cam);
}

// ---------------------------------------------------------------------------
// Per-pixel loop (geodesic_kernel.cu:462-510).  Outputs:
//   hdr  : mean over samples of the raw radiance (before exposure)
//   ldr  : mean over samples of aces(hdr*exposure)  (reference accumulation)
//   rgba : to_u8(linear_to_srgb(ldr)), alpha 255
//   cls  : class of sample 0;  steps: sum of RK4 steps over the samples
// ---------------------------------------------------------------------------
template <class T> struct PixelOut {
    double hdr[3];
    double ldr[3];
    std::uint8_t rgba[4];
    std::uint8_t cls;
    std::uint32_t steps;
    double inv_drift[3];  // max |inv1 - inv0| over samples
};

template <class T>
inline PixelOut<T> render_pixel(const Uni<T>& u, const CamFrame<T>& f, std::uint32_t gx, std::uint32_t gy,
                                const float* halton_xy) {
    PixelOut<T> p{};
    std::uint32_t spp = u.supersample;
    spp = spp < 1u ? 1u : (spp > 1024u ? 1024u : spp);
    // realtime backends' sampler (render/metal/shaders/geodesic_kernel.metal:416-434):
    // ss x ss grid at ((sx + 0.5) / ss, (sy + 0.5) / ss), mean radiance, then tone map
    const std::uint32_t ss = spp < 4u ? spp : 4u;
    if (u.ext.grid) spp = ss * ss;
    const float sub_w = 1.0f / float(ss);
    V3<T> ldr = v3(T(0.0), T(0.0), T(0.0));
    V3<T> hdr = v3(T(0.0), T(0.0), T(0.0));
    for (std::uint32_t s = 0; s < spp; ++s) {
        T ox = T(0.5), oy = T(0.5);
        if (u.ext.grid) {
            ox = T(double((float(s % ss) + 0.5f) * sub_w));
            oy = T(double((float(s / ss) + 0.5f) * sub_w));
        } else if (spp > 1u) {
            ox = T(double(halton_xy[2 * s]));
            oy = T(double(halton_xy[2 * s + 1]));
        }
        if (val(u.ext.jitter_x) != 0.0 || val(u.ext.jitter_y) != 0.0) {
            ox = ox + u.ext.jitter_x;
            oy = oy + u.ext.jitter_y;
        }
        const RayOut<T> r = trace_ray(u, f, gx, gy, ox, oy);
        if (s == 0) p.cls = std::uint8_t(r.cls);
        p.steps += r.steps;
        for (int k = 0; k < 3; ++k) {
            const double d = std::fabs(r.inv1[k] - r.inv0[k]);
            if (d > p.inv_drift[k] || std::isnan(d)) p.inv_drift[k] = d;
        }
        hdr = hdr + r.rgb;
        ldr = ldr + aces(r.rgb * u.exposure);
    }
    if (spp > 1u) {
        const T inv = T(1.0) / T(double(spp));
        ldr = ldr * inv;
        hdr = hdr * inv;
    }
    if (u.ext.grid) ldr = aces(hdr * u.exposure);
    p.hdr[0] = val(hdr.x); p.hdr[1] = val(hdr.y); p.hdr[2] = val(hdr.z);
    p.ldr[0] = val(ldr.x); p.ldr[1] = val(ldr.y); p.ldr[2] = val(ldr.z);
    p.rgba[0] = to_u8(ldr.x);
    p.rgba[1] = to_u8(ldr.y);
    p.rgba[2] = to_u8(ldr.z);
    p.rgba[3] = 255;
    return p;
}

inline void halton_table(std::uint32_t spp, float* out) {
    for (std::uint32_t s = 0; s < spp; ++s) {
        out[2 * s] = halton(s + 1u, 2u);
        out[2 * s + 1] = halton(s + 1u, 3u);
    }
}

// Row/pixel work stealing over std::thread workers (cli/main.cpp:739-740,861-872).
template <class F> inline void parallel_for(std::int64_t n, int threads, F fn) {
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<std::int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const std::int64_t i = next.fetch_add(1);
            if (i >= n) break;
            fn(i);
        }
    };
    if (threads == 1) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work);
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Post passes (render/vulkan/shaders/bloom.hlsl:8-77, blit.hlsl:43-58), on
// float RGB HDR input:  extract (2x2 Karis + soft knee) at floor(W/2) x floor(H/2),
// 7-tap separable Gaussian H then V with clamp-to-edge, composite
// (hdr + bloom*strength)*exposure -> ACES -> sRGB u8.
// ---------------------------------------------------------------------------
// IEEE binary16 round-to-nearest-even storage emulation (RGBA16F targets).
inline float to_half_and_back(float x) {
    std::uint32_t b;
    std::memcpy(&b, &x, 4);
    const std::uint32_t sign = b & 0x80000000u;
    std::uint32_t ab = b & 0x7FFFFFFFu;
    if (ab >= 0x7F800000u) return x;                      // inf / nan unchanged
    if (ab >= 0x477FF000u) {                              // rounds to >= 65520 -> inf
        const std::uint32_t o = sign | 0x7F800000u;
        float r; std::memcpy(&r, &o, 4); return r;
    }
    float fa; std::memcpy(&fa, &ab, 4);
    float q;
    if (ab < 0x38800000u) {                               // half subnormal range: quantum 2^-24
        const float quantum = 5.9604644775390625e-08f;
        q = std::nearbyint(fa / quantum) * quantum;
    } else {                                              // normal: keep 10 mantissa bits, RNE
        const std::uint32_t lsb = (ab >> 13) & 1u;
        const std::uint32_t rounded = (ab + 0x0FFFu + lsb) & ~0x1FFFu;
        std::memcpy(&q, &rounded, 4);
    }
    std::uint32_t qb; std::memcpy(&qb, &q, 4);
    qb |= sign;
    float r; std::memcpy(&r, &qb, 4);
    return r;
}

inline const float* bloom_weights() {  // bloom.hlsl:8-11
    static const float w[7] = {0.071303f, 0.131514f, 0.189879f, 0.214607f, 0.189879f, 0.131514f, 0.071303f};
    return w;
}
inline float lum3(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }

// bloom.hlsl:19-47.  hdr: W*H*3 float (already in storage precision).
inline void bloom_extract(const float* hdr, int W, int H, float threshold, float* out, bool half_store) {
    const int Wb = W / 2, Hb = H / 2;
    for (int y = 0; y < Hb; ++y)
        for (int x = 0; x < Wb; ++x) {
            const float* t[4] = {hdr + 3 * ((2 * y) * W + 2 * x), hdr + 3 * ((2 * y) * W + 2 * x + 1),
                                 hdr + 3 * ((2 * y + 1) * W + 2 * x), hdr + 3 * ((2 * y + 1) * W + 2 * x + 1)};
            float w[4], ws = 0.0f;
            for (int k = 0; k < 4; ++k) {
                w[k] = 1.0f / (1.0f + lum3(t[k][0], t[k][1], t[k][2]));
            }
            ws = ((w[0] + w[1]) + w[2]) + w[3];
            const float den = std::fmax(ws, 1e-6f);
            float avg[3];
            for (int c = 0; c < 3; ++c)
                avg[c] = (((t[0][c] * w[0] + t[1][c] * w[1]) + t[2][c] * w[2]) + t[3][c] * w[3]) / den;
            const float lum = lum3(avg[0], avg[1], avg[2]);
            const float knee = (lum * lum) / std::fmax(lum + threshold, 1e-4f);
            const float k = (lum > 1e-4f) ? (knee / lum) : 0.0f;
            for (int c = 0; c < 3; ++c) {
                const float v = avg[c] * k;
                out[3 * (y * Wb + x) + c] = half_store ? to_half_and_back(v) : v;
            }
        }
}
// bloom.hlsl:49-77: 7-tap, clamp-to-edge. horizontal = true for blur_h.
inline void bloom_blur(const float* in, int Wb, int Hb, bool horizontal, float* out, bool half_store) {
    const float* w = bloom_weights();
    for (int y = 0; y < Hb; ++y)
        for (int x = 0; x < Wb; ++x) {
            float s[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < 7; ++i) {
                int px = x, py = y;
                if (horizontal) px = std::min(std::max(x + i - 3, 0), Wb - 1);
                else py = std::min(std::max(y + i - 3, 0), Hb - 1);
                const float* v = in + 3 * (py * Wb + px);
                for (int c = 0; c < 3; ++c) s[c] = s[c] + v[c] * w[i];
            }
            for (int c = 0; c < 3; ++c) out[3 * (y * Wb + x) + c] = half_store ? to_half_and_back(s[c]) : s[c];
        }
}
// blit.hlsl:43-58: c = hdr + bilinear(bloom)*strength; c *= exposure; ACES; sRGB; u8.
// The half-res bloom is sampled at the full-res pixel centre, clamp-to-edge.
inline void bloom_composite(const float* hdr, const float* bloom, int W, int H, float strength, float exposure,
                            std::uint8_t* rgba) {
    const int Wb = W / 2, Hb = H / 2;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float px = (float(x) + 0.5f) * float(Wb) / float(W) - 0.5f;
            const float py = (float(y) + 0.5f) * float(Hb) / float(H) - 0.5f;
            const float fx0 = std::floor(px), fy0 = std::floor(py);
            const float fx = px - fx0, fy = py - fy0;
            const int x0 = std::min(std::max(int(fx0), 0), Wb - 1), x1 = std::min(std::max(int(fx0) + 1, 0), Wb - 1);
            const int y0 = std::min(std::max(int(fy0), 0), Hb - 1), y1 = std::min(std::max(int(fy0) + 1, 0), Hb - 1);
            V3<float> c;
            float b[3];
            for (int k = 0; k < 3; ++k) {
                const float top = bloom[3 * (y0 * Wb + x0) + k] + fx * (bloom[3 * (y0 * Wb + x1) + k] - bloom[3 * (y0 * Wb + x0) + k]);
                const float bot = bloom[3 * (y1 * Wb + x0) + k] + fx * (bloom[3 * (y1 * Wb + x1) + k] - bloom[3 * (y1 * Wb + x0) + k]);
                b[k] = top + fy * (bot - top);
            }
            const float* h = hdr + 3 * (y * W + x);
            c.x = (h[0] + b[0] * strength) * exposure;
            c.y = (h[1] + b[1] * strength) * exposure;
            c.z = (h[2] + b[2] * strength) * exposure;
            const V3<float> t = aces(c);
            std::uint8_t* o = rgba + 4 * (y * W + x);
            o[0] = to_u8(t.x);
            o[1] = to_u8(t.y);
            o[2] = to_u8(t.z);
            o[3] = 255;
        }
}

// ---------------------------------------------------------------------------
// The CPU driver's diagnostics (cli/main.cpp:1269-1621): Kerr geometry
// scalars, the circular photon orbit trail and the benchmark ray fan.  The
// product runs the orbit and the fan on the device; these are its checker.
// ---------------------------------------------------------------------------
template <class T> inline T kerr_ergo_outer(T M, T a, T theta) {  // kerr.hpp:53-57
    const T c = xcos(theta);
    const T d = M * M - a * a * c * c;
    return M + xsqrt(d > T(0.0) ? d : T(0.0));
}

// cli/main.cpp:1415-1476, in output order: mass, spin, spin/mass, r+, r-,
// ergosphere at the pole and at the equator, photon orbits (pro, retro),
// ISCO (pro, retro).  Returns 2 (usage error) for M <= 0 or a outside [0, M].
template <class T> inline int kerr_geometry(T M, T a, T* out) {
    if (M <= T(0.0) || a < T(0.0) || a > M) return 2;
    const T half_pi = T(0.5) * T(3.14159265358979323846);
    out[0] = M;
    out[1] = a;
    out[2] = a / M;
    out[3] = kerr_r_plus(M, a);
    out[4] = kerr_r_minus(M, a);
    out[5] = kerr_ergo_outer(M, a, T(0.0));
    out[6] = kerr_ergo_outer(M, a, half_pi);
    out[7] = kerr_photon_pro(M, a);
    out[8] = kerr_photon_ret(M, a);
    out[9] = kerr_isco(M, a, true);
    out[10] = kerr_isco(M, a, false);
    return 0;
}

struct OrbitRow {
    std::int32_t step;  // -1: the sentinel row after a ray left the orbit (non-finite or inside the cut)
    double t, r, theta, phi;
};

// cli/main.cpp:1279-1410: one equatorial circular null orbit at r_init,
// integrated for `orbits` periods with step h.  Rows: the initial state, then
// every sub_every = max(1, int(period / (200 h))) steps and the last step, or a
// -1 sentinel row where the ray leaves (r non-finite or below 1.001 r+).
// Schwarzschild (a = 0): the 8-component geodesic, u^phi = sqrt(1 - 2M/r)/r.
// Kerr: the Hamiltonian form with E = 1, L_z = (r sqrt(Delta) - 2 M a)/(r - 2M),
// Q = 0; the period comes from d(phi)/d(lambda) at the start.  Returns 0, or 2
// for the argument errors the driver rejects.
template <class T>
inline int photon_orbit(T M, T a, T r_init, T orbits, T h, std::vector<OrbitRow>& rows) {
    rows.clear();
    const T pi = T(3.14159265358979323846);
    if (M <= T(0.0) || a < T(0.0) || a > M) return 2;
    const T rs = T(2.0) * M;
    const T r_plus = (a > T(0.0)) ? kerr_r_plus(M, a) : rs;
    if (r_init <= r_plus * T(1.001) || orbits <= T(0.0) || h <= T(0.0)) return 2;
    auto emit = [&](int step, T t, T r, T th, T ph) { rows.push_back({step, double(t), double(r), double(th), double(ph)}); };
    if (a <= T(0.0)) {
        const T f = T(1.0) - rs / r_init;
        SState<T> y{T(0.0), r_init, T(0.5) * pi, T(0.0), T(1.0), T(0.0), T(0.0), xsqrt(f) / r_init};
        const T period = T(2.0) * pi / y.uphi;
        const int total = int(orbits * period / h);
        const int every = std::max(1, int(period / (h * T(200.0))));
        emit(0, y.t, y.r, y.theta, y.phi);
        for (int step = 1; step <= total; ++step) {
            y = schw_rk4(y, h, rs);
            if (!xisfinite(y.r) || y.r < rs * T(1.001)) {
                emit(-1, y.t, y.r, y.theta, y.phi);
                break;
            }
            if (step % every == 0 || step == total) emit(step, y.t, y.r, y.theta, y.phi);
        }
        return 0;
    }
    const T Delta = r_init * r_init - T(2.0) * M * r_init + a * a;
    if (Delta <= T(0.0)) return 2;
    const T b_pro = (r_init * xsqrt(Delta) - T(2.0) * M * a) / (r_init - T(2.0) * M);
    const KConst<T> c{T(1.0), b_pro, T(0.0), a, M};
    KState<T> y{T(0.0), r_init, T(0.5) * pi, T(0.0), T(0.0), T(0.0)};
    const KState<T> d0 = kerr_rhs(y, c);
    if (!xisfinite(d0.phi) || d0.phi <= T(0.0)) return 2;
    const T period = T(2.0) * pi / d0.phi;
    const int total = int(orbits * period / h);
    const int every = std::max(1, int(period / (h * T(200.0))));
    emit(0, y.t, y.r, y.theta, y.phi);
    for (int step = 1; step <= total; ++step) {
        y = kerr_rk4(y, h, c);
        if (!xisfinite(y.r) || y.r < r_plus * T(1.001)) {
            emit(-1, y.t, y.r, y.theta, y.phi);
            break;
        }
        if (step % every == 0 || step == total) emit(step, y.t, y.r, y.theta, y.phi);
    }
    return 0;
}

// cli/main.cpp:1492-1590: the benchmark's equatorial ray fan (M = 1, h = 0.1).
// Ray i has impact parameter b_i = b_min + (b_max - b_min)(i + 1/2)/n and
// starts at x = max(50, 2 max|b|) on the line y = b moving in -x.  Each ray
// steps until max_steps, a non-finite r or phi, or r below the cut (1.01 r_s
// Schwarzschild, 1.02 r+ Kerr).  Kerr rays take u^t from the null condition
// (future root) and become (E, L_z, Q = 0) and p_r = Sigma u^r / Delta; a ray
// whose quadratic has no real root is skipped (0 steps).  per_ray (may be null)
// gets each ray's step count; the return value is the total.
template <class T>
inline long long bench_fan(bool kerr, T spin, int n_rays, int max_steps, T b_min, T b_max, std::uint32_t* per_ray) {
    const T pi = T(3.14159265358979323846);
    const T M = T(1.0), rs = T(2.0) * M, h = T(0.1);
    const T a = kerr ? spin * M : T(0.0);
    const T bmin = b_min * M, bmax = b_max * M;
    const T ab_min = bmin < T(0.0) ? -bmin : bmin, ab_max = bmax < T(0.0) ? -bmax : bmax;
    const T x0 = std::max(T(50.0) * M, T(2.0) * std::max(ab_min, ab_max));
    long long total = 0;
    for (int i = 0; i < n_rays; ++i) {
        const T b = bmin + (bmax - bmin) * (T(double(i)) + T(0.5)) / T(double(n_rays));
        const T r0 = xsqrt(x0 * x0 + b * b);
        const T phi0 = xatan2(b, x0);
        const T ur = -xcos(phi0);
        const T uphi = xsin(phi0) / r0;
        int n = 0;
        if (!kerr) {
            const T cut = T(1.01) * rs;
            const T f = T(1.0) - rs / r0;
            const T ut = xsqrt(((ur * ur) / f + r0 * r0 * uphi * uphi) / f);
            SState<T> y{T(0.0), r0, T(0.5) * pi, phi0, ut, ur, T(0.0), uphi};
            for (int step = 0; step < max_steps; ++step) {
                if (!xisfinite(y.r) || !xisfinite(y.phi)) break;
                if (y.r < cut) break;
                y = schw_rk4(y, h, rs);
                ++n;
            }
        } else {
            const T cut = T(1.02) * kerr_r_plus(M, a);
            const T Delta = r0 * r0 - T(2.0) * M * r0 + a * a;
            const T g_tt = -(T(1.0) - T(2.0) * M / r0);
            const T g_tp = -T(2.0) * M * a / r0;
            const T g_rr = r0 * r0 / Delta;
            const T g_pp = ((r0 * r0 + a * a) * (r0 * r0 + a * a) - a * a * Delta) / (r0 * r0);
            const T qa = g_tt, qb = T(2.0) * g_tp * uphi, qc = g_rr * ur * ur + g_pp * uphi * uphi;
            const T disc = qb * qb - T(4.0) * qa * qc;
            if (disc >= T(0.0)) {
                const T ut = (-qb - xsqrt(disc)) / (T(2.0) * qa);
                const KConst<T> c{-(g_tt * ut + g_tp * uphi), g_tp * ut + g_pp * uphi, T(0.0), a, M};
                KState<T> y{T(0.0), r0, T(0.5) * pi, phi0, T(0.0), T(0.0)};
                // kerr_hamilton.h:192-200: p_r = Sigma u^r / Delta, p_theta = Sigma u^theta (= 0)
                const T ct = xcos(y.theta);
                const T Sig = r0 * r0 + a * a * ct * ct;
                const T Del = r0 * r0 - T(2.0) * M * r0 + a * a;
                y.p_r = Sig * ur / Del;
                y.p_theta = Sig * T(0.0);
                for (int step = 0; step < max_steps; ++step) {
                    if (!xisfinite(y.r) || !xisfinite(y.phi)) break;
                    if (y.r < cut) break;
                    y = kerr_rk4(y, h, c);
                    ++n;
                }
            }
        }
        if (per_ray) per_ray[i] = std::uint32_t(n);
        total += n;
    }
    return total;
}

}  // namespace orc
