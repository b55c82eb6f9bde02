// singularity_amd/csrc/physics.hpp — device-side physics of the trace kernels,
// templated on the arithmetic type R (float or double).
//
// The algorithm is the reference's (shared_shader/kerr_hamilton.h:97-178,
// shared_shader/geodesic_math.h:101-134, shared_shader/disc_intersection.h:23-55,
// render/cuda/kernels/geodesic_kernel.cu:77-403); the arithmetic is arranged
// for the CDNA4 vector ALU:
//   * the integrated time coordinate t is dropped (no output depends on it);
//   * each RHS evaluates one reciprocal (1/(Sigma Delta sin^2)) instead of
//     13 divides, and per-ray products of E and L_z are formed once per ray;
//   * sin/cos of theta for the pole damping are reused by the first RK4 stage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "scene_const.hpp"

namespace sing {

// ---- scalar helpers ---------------------------------------------------------
// sin and cos of the polar angle on the step path: Cody-Waite reduction by
// pi/2 (two-term constant, exact with FMA for the |x| < 2^20 this path sees)
// to r in [-pi/4, pi/4], Taylor polynomials of sin r and cos r (coefficients
// (-1)^n/n!; truncation below half an ulp on that interval), quadrant swap.
// Branch-free and short-lived in registers, unlike a general-argument sincos.
__device__ __forceinline__ void dsincos(float x, float* s, float* c) {
    const float k = rintf(x * 0.636619772367581343f);
    float r = fmaf(-k, 1.5707963705062866f, x);
    r = fmaf(-k, -4.371138828673793e-08f, r);
    const float r2 = r * r;
    float ps = fmaf(r2, 2.7557319e-06f, -1.9841270e-04f);
    ps = fmaf(r2, ps, 8.3333338e-03f);
    ps = fmaf(r2, ps, -1.6666667e-01f);
    const float sr = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.7557320e-07f, 2.4801588e-05f);
    pc = fmaf(r2, pc, -1.3888889e-03f);
    pc = fmaf(r2, pc, 4.1666668e-02f);
    pc = fmaf(r2, pc, -0.5f);
    const float cr = fmaf(r2, pc, 1.0f);
    const int q = int(k);
    float ss = (q & 1) ? cr : sr;
    float cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
__device__ __forceinline__ void dsincos(double x, double* s, double* c) {
    const double k = rint(x * 0.63661977236758134);
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    const double r2 = r * r;
    double ps = fma(r2, 2.8114572543455206e-15, -7.647163731819816e-13);
    ps = fma(r2, ps, 1.6059043836821613e-10);
    ps = fma(r2, ps, -2.505210838544172e-08);
    ps = fma(r2, ps, 2.7557319223985893e-06);
    ps = fma(r2, ps, -0.0001984126984126984);
    ps = fma(r2, ps, 0.008333333333333333);
    ps = fma(r2, ps, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = fma(r2, 4.779477332387385e-14, -1.1470745597729725e-11);
    pc = fma(r2, pc, 2.08767569878681e-09);
    pc = fma(r2, pc, -2.755731922398589e-07);
    pc = fma(r2, pc, 2.48015873015873e-05);
    pc = fma(r2, pc, -0.001388888888888889);
    pc = fma(r2, pc, 0.041666666666666664);
    pc = fma(r2, pc, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int q = int(k);
    double ss = (q & 1) ? cr : sr;
    double cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
// sin and cos of th0 + d from s0 = sin th0 and c0 = cos th0 by the angle-
// addition formulas.  An RK4 stage moves theta by a small d (|d| > 0.125 in
// about 3 of 10^4 stages of the a = 0.94 headline scene), so for |d| <= 0.125
// sin d and cos d come from short Taylor polynomials and the stage needs no
// reduction or quadrant selection; a larger |d| takes dsincos.  Truncation on
// |d| <= 0.125: FP32 sin to d^5 (next term d^7/7! < 1e-10), cos to d^4 (next
// d^6/6! < 5.3e-9, 0.09 ulp); FP64 sin to d^11 (next d^13/13! < 3e-22), cos
// to d^10 (next d^12/12! < 3e-20).
__device__ __forceinline__ void sincos_shift(float th0, float d, float s0, float c0, float* s, float* c) {
    if (__builtin_expect(fabsf(d) <= 0.125f, 1)) {
        const float d2 = d * d;
        const float sd = fmaf(d * d2, fmaf(d2, 8.3333333e-03f, -1.6666667e-01f), d);
        const float cd = fmaf(d2, fmaf(d2, 4.1666668e-02f, -0.5f), 1.0f);
        *s = fmaf(s0, cd, c0 * sd);
        *c = fmaf(c0, cd, -(s0 * sd));
    } else {
        dsincos(th0 + d, s, c);
    }
}
__device__ __forceinline__ void sincos_shift(double th0, double d, double s0, double c0, double* s, double* c) {
    if (__builtin_expect(fabs(d) <= 0.125, 1)) {
        // sums of coefficient x power, highest first: every FMA accumulates
        // into the running sum, so a coefficient is only ever a multiplicand.
        // In Horner's form it is the addend, which the compiler turned into a
        // 64-bit register copy plus a two-address FMA per term in the
        // register-bound fixed-step FP64 passes (this rotation runs three
        // times per step there).  The full evaluation above keeps Horner: it
        // runs once per 8 steps in those passes, and in the error-controlled
        // FP64 pass, which evaluates it every stage, the power form was
        // slower.  Same accuracy either way (<= 0.51 ulp here).
        const double d2 = d * d, d4 = d2 * d2, d6 = d4 * d2, d8 = d4 * d4;
        double ps = 2.505210838544172e-08 * d8;
        ps = fma(-2.755731922398589e-06, d6, ps);
        ps = fma(1.984126984126984e-04, d4, ps);
        ps = fma(-8.333333333333333e-03, d2, ps);
        ps = ps + 1.6666666666666666e-01;
        const double sd = fma(-(d * d2), ps, d);
        double pc = -2.755731922398589e-07 * d8;
        pc = fma(2.48015873015873e-05, d6, pc);
        pc = fma(-1.388888888888889e-03, d4, pc);
        pc = fma(4.1666666666666664e-02, d2, pc);
        pc = pc - 0.5;
        const double cd = fma(d2, pc, 1.0);
        *s = fma(s0, cd, c0 * sd);
        *c = fma(c0, cd, -(s0 * sd));
    } else {
        dsincos(th0 + d, s, c);
    }
}
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double dpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float dcos(float x) { return cosf(x); }
__device__ __forceinline__ double dcos(double x) { return cos(x); }
__device__ __forceinline__ float dsin(float x) { return sinf(x); }
__device__ __forceinline__ double dsin(double x) { return sin(x); }
__device__ __forceinline__ float dacos(float x) { return acosf(x); }
__device__ __forceinline__ double dacos(double x) { return acos(x); }
__device__ __forceinline__ float datan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double datan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float dmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float dmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double dmin(double a, double b) { return fmin(a, b); }
// Reciprocal on the RHS hot path: v_rcp_f32 (1 ulp) in FP32, exact division in FP64.
__device__ __forceinline__ float drcp(float x) { return __builtin_amdgcn_rcpf(x); }
// Double reciprocal: the hardware estimate refined by two Newton steps
// (e = 1 - x r, r += r e; the error squares each step), 5 instructions
// instead of the ~10 of an IEEE division.  Used for the Kerr RHS's
// 1/(Sigma Delta sin^2) and the Schwarzschild 1/r, 1/f, 1/sin, whose
// arguments stay finite and far from zero and overflow on the path.
__device__ __forceinline__ double drcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
template <class R> __device__ __forceinline__ R dclamp(R x, R lo, R hi) { return dmin(dmax(x, lo), hi); }
// max(x, lo) for x = sin^2 theta >= 0 and lo > 0.  Non-negative floats order
// like their bit patterns, so in FP32 this is one integer max (fmaxf would
// first canonicalise its operand).
__device__ __forceinline__ float pos_max(float x, float lo) {
    return __uint_as_float(max(__float_as_uint(x), __float_as_uint(lo)));
}
__device__ __forceinline__ double pos_max(double x, double lo) { return fmax(x, lo); }
// min(x, y) of two non-negative values (one integer min in FP32, as pos_max).
__device__ __forceinline__ float pos_min(float x, float y) {
    return __uint_as_float(min(__float_as_uint(x), __float_as_uint(y)));
}
__device__ __forceinline__ double pos_min(double x, double y) { return fmin(x, y); }
__device__ __forceinline__ float dabs(float x) { return fabsf(x); }
__device__ __forceinline__ double dabs(double x) { return fabs(x); }

// ---- Kerr Hamiltonian RHS (kerr_hamilton.h:97-155) ----------------------------
// State (r, theta, phi, p_r, p_theta); st/ct = sin/cos(theta).
template <class R> struct KerrK {
    R E, L, L2, aE, a2E2, aL, twoE, c2aEL;  // E, L_z, L_z^2, a E, a^2 E^2, a L_z, 2 E, -2 a E L_z
};
template <class R> __device__ __forceinline__ KerrK<R> kerr_k(const SceneConst<R>& sc, R E, R L) {
    KerrK<R> k;
    k.E = E;
    k.L = L;
    k.L2 = L * L;
    k.aE = sc.a * E;
    k.a2E2 = k.aE * k.aE;
    k.aL = sc.a * L;
    k.twoE = E + E;
    k.c2aEL = R(-2) * k.aE * L;
    return k;
}

// The reference's Hamiltonian is 2H = B/(Sigma Delta) + L^2/(Sigma s2) +
// (Delta p_r^2 + p_th^2)/Sigma with B = -A E^2 + 4 M r a E L - a^2 L^2,
// A = (r^2+a^2)^2 - Delta a^2 sin^2, s2 = max(sin^2, 1e-10)
// (kerr_hamilton.h:67-94, 121-152).  Since r^2 + a^2 - Delta = 2 M r, the same
// function is 2H = F / Sigma with Carter's separated numerator
//   F = Delta p_r^2 + p_th^2 - P^2/Delta + L^2/s2 - 2 a E L + a^2 E^2 sin^2,
//   P = (r^2 + a^2) E - a L,
// and Hamilton's equations (dq/dl = dH/dp, dp/dl = -dH/dq, H = F/(2 Sigma)):
//   dr/dl    = Delta p_r / Sigma,          dth/dl = p_th / Sigma,
//   dphi/dl  = (a P/Delta + L/s2 - a E) / Sigma,
//   dp_r/dl  = (2 r E P/Delta - (r - M)(p_r^2 + P^2/Delta^2) + r F/Sigma) / Sigma,
//   dp_th/dl = sin cos ((L^2/s2^2 - a^2 E^2) - a^2 F/Sigma) / Sigma
// — the reference's partials of the same 2H, clamp included (its dphi/dl
// numerator has sin^2 where this has s2: the two differ only below
// |sin| = 1e-5, by 1e-10 relative).  44 operations and one reciprocal per
// evaluation instead of the direct form's ~57 and three.
// T_OUT: also dt/dl = (A E - 2 M a r L_z)/(Sigma Delta) = ((r^2+a^2) P/Delta +
// a L - a^2 E sin^2)/Sigma into *dt (the renderer does not track t; the
// photon-orbit diagnostic does).
template <class R, bool T_OUT = false>
__device__ __forceinline__ void kerr_rhs_sc(R r, R pr, R pth, R st, R ct, const SceneConst<R>& sc, const KerrK<R>& k,
                                            R& dr, R& dth, R& dph, R& dpr, R& dpth, R* dt = nullptr) {
    const R a2 = sc.a2;
    const R s2 = st * st;
    const R safe = pos_max(s2, R(1e-10));
    const R r2a2 = r * r + a2;
    const R Sig = r2a2 - a2 * s2;
    const R Del = r * (r - sc.twoM) + a2;
    const R P = r2a2 * k.E - k.aL;
    // one reciprocal: q = 1/(Sigma Delta s2), then 1/Sigma, 1/Delta and 1/s2
    // by multiplying back
    const R SD = Sig * Del;
    const R q = drcp(SD * safe);
    const R iS = q * (Del * safe);
    const R iD = q * (Sig * safe);
    const R u = q * SD;
    const R PiD = P * iD;
    const R pr2 = pr * pr;
    const R F = ((Del * pr2 + pth * pth) - P * PiD) + (k.L2 * u + k.c2aEL) + k.a2E2 * s2;
    dr = Del * pr * iS;
    dth = pth * iS;
    dph = (sc.a * PiD + (k.L * u - k.aE)) * iS;
    if constexpr (T_OUT) *dt = (r2a2 * PiD + (k.aL - sc.a * k.aE * s2)) * iS;
    const R FiS = F * iS;
    const R rmM = r - sc.M;
    dpr = (((k.twoE * r) * PiD - rmM * (PiD * PiD + pr2)) + r * FiS) * iS;
    dpth = ((st * ct) * iS) * ((k.L2 * (u * u) - k.a2E2) - a2 * FiS);
}

// The reference's own grouping of the same RHS (kerr_hamilton.h:121-152:
// 2H = B/(Sigma Delta) + L^2/(Sigma s2) + (Delta p_r^2 + p_th^2)/Sigma and its
// quotient-rule partials), with 1/Sigma, 1/Delta, 1/s2 as three reciprocals.
// Used by the CPU driver's diagnostics (orbit, ray fan), whose fixed-step
// rays near capture are chaotic: step counts there depend on rounding, and
// this grouping reproduces the driver's totals.
template <class R> struct KerrKDirect {
    R E, EE, L, L2, cB1, cB0, cphi, mhcB1;  // E, E^2, L_z, L_z^2, 4 M a E L_z, a^2 L_z^2, 2 M a E, -2 M a E L_z
};
template <class R> __device__ __forceinline__ KerrKDirect<R> kerr_k_direct(const SceneConst<R>& sc, R E, R L) {
    KerrKDirect<R> k;
    k.E = E;
    k.EE = E * E;
    k.L = L;
    k.L2 = L * L;
    k.cB1 = R(2) * sc.twoM * sc.a * E * L;
    k.cB0 = sc.a2 * k.L2;
    k.cphi = sc.twoM * sc.a * E;
    k.mhcB1 = R(-0.5) * k.cB1;
    return k;
}
template <class R, bool T_OUT = false>
__device__ __forceinline__ void kerr_rhs_sc(R r, R pr, R pth, R st, R ct, const SceneConst<R>& sc,
                                            const KerrKDirect<R>& k, R& dr, R& dth, R& dph, R& dpr, R& dpth,
                                            R* dt = nullptr) {
    // the -1/2 of dp/dl = -(1/2) d(2H)/dq folded into each partial, Sigma =
    // r^2 + a^2 - a^2 sin^2, Delta/(Sigma Delta) = 1/Sigma, the d/dtheta terms
    // sharing sin cos (d Sigma/d theta = -2 a^2 sin cos, dA/d theta = -2 a^2
    // Delta sin cos)
    const R a2 = sc.a2;
    const R s2 = st * st;
    const R safe = pos_max(s2, R(1e-10));
    const R r2 = r * r;
    const R as2 = a2 * s2;
    const R r2a2 = r2 + a2;
    const R Sig = r2a2 - as2;
    const R Del = r2 - sc.twoM * r + a2;
    const R A = r2a2 * r2a2 - Del * as2;
    const R B = (k.cB1 * r - k.cB0) - A * k.EE;
    const R iS = drcp(Sig);
    const R iD = drcp(Del);
    const R iss = drcp(safe);
    const R iSD = iS * iD;
    const R iSiss = iS * iss;
    dr = Del * pr * iS;
    dth = pth * iS;
    dph = (k.cphi * r * s2 + (Del - as2) * k.L) * (iD * iSiss);
    if constexpr (T_OUT) *dt = (A * k.E - sc.twoM * sc.a * r * k.L) * iSD;
    const R rmM = r - sc.M;                                       // (1/2) d Delta/dr
    const R hBr = ((r + r) * r2a2 - rmM * as2) * k.EE + k.mhcB1;  // -(1/2) dB/dr
    const R BiSD = B * iSD;
    const R Xh = r * Del + Sig * rmM;                             // (1/2) d(Sigma Delta)/dr
    const R pr2 = pr * pr;
    const R Dp2 = Del * pr2 + pth * pth;
    dpr = hBr * iSD + BiSD * Xh * iSD + k.L2 * r * (iS * iSiss) - rmM * pr2 * iS + Dp2 * r * (iS * iS);
    const R Y = (k.EE + BiSD) + Dp2 * iS;
    dpth = (st * ct) * (k.L2 * (Sig - a2 * safe) * (iSiss * iSiss) - (a2 * iS) * Y);
}

// Pole damping factor clamp(5|sin theta|, 0.05, 1) (kerr_hamilton.h:166-169).
template <class R> __device__ __forceinline__ R kerr_pole_scale(R st) {
    const R ps = (st < R(0) ? -st : st) * R(5);
    return (ps > R(1)) ? R(1) : ((ps < R(0.05)) ? R(0.05) : ps);
}
// FP32: |sin| * 5 with the abs folded into the multiply, clamped by one med3.
template <> __device__ __forceinline__ float kerr_pole_scale<float>(float st) {
    return __builtin_amdgcn_fmed3f(fabsf(st) * 5.0f, 0.05f, 1.0f);
}

// Classical RK4 step of size h (kerr_hamilton.h:171-177); (st, ct) = sin/cos of
// the starting theta, reused by the first stage.  T_OUT: t advances too (*t).
// SHIFT: stages 2-4 rotate (sin, cos) of the starting theta by the stage's
// theta increment (sincos_shift) instead of evaluating sincos afresh.
template <class R, bool T_OUT = false, bool SHIFT = true, class K = KerrK<R>>
__device__ __forceinline__ void kerr_rk4_raw(R& r, R& th, R& ph, R& pr, R& pth, R h, R st, R ct,
                                             const SceneConst<R>& sc, const K& k, R* t = nullptr) {
    const R hh = R(0.5) * h;
    const R st0 = st, ct0 = ct;
    R k_r, k_th, k_ph, k_pr, k_pth, k_t = R(0);
    kerr_rhs_sc<R, T_OUT>(r, pr, pth, st, ct, sc, k, k_r, k_th, k_ph, k_pr, k_pth, &k_t);
    R a_r = k_r, a_th = k_th, a_ph = k_ph, a_pr = k_pr, a_pth = k_pth, a_t = k_t;
    // the stages need theta only through sin/cos: rotate (st0, ct0) by the
    // stage's theta increment
    R y_r = r + hh * k_r, y_pr = pr + hh * k_pr, y_pth = pth + hh * k_pth;
    if constexpr (SHIFT) sincos_shift(th, hh * k_th, st0, ct0, &st, &ct);
    else dsincos(th + hh * k_th, &st, &ct);
    kerr_rhs_sc<R, T_OUT>(y_r, y_pr, y_pth, st, ct, sc, k, k_r, k_th, k_ph, k_pr, k_pth, &k_t);
    a_r += R(2) * k_r; a_th += R(2) * k_th; a_ph += R(2) * k_ph; a_pr += R(2) * k_pr; a_pth += R(2) * k_pth;
    if constexpr (T_OUT) a_t += R(2) * k_t;
    y_r = r + hh * k_r; y_pr = pr + hh * k_pr; y_pth = pth + hh * k_pth;
    if constexpr (SHIFT) sincos_shift(th, hh * k_th, st0, ct0, &st, &ct);
    else dsincos(th + hh * k_th, &st, &ct);
    kerr_rhs_sc<R, T_OUT>(y_r, y_pr, y_pth, st, ct, sc, k, k_r, k_th, k_ph, k_pr, k_pth, &k_t);
    a_r += R(2) * k_r; a_th += R(2) * k_th; a_ph += R(2) * k_ph; a_pr += R(2) * k_pr; a_pth += R(2) * k_pth;
    if constexpr (T_OUT) a_t += R(2) * k_t;
    y_r = r + h * k_r; y_pr = pr + h * k_pr; y_pth = pth + h * k_pth;
    if constexpr (SHIFT) sincos_shift(th, h * k_th, st0, ct0, &st, &ct);
    else dsincos(th + h * k_th, &st, &ct);
    kerr_rhs_sc<R, T_OUT>(y_r, y_pr, y_pth, st, ct, sc, k, k_r, k_th, k_ph, k_pr, k_pth, &k_t);
    a_r += k_r; a_th += k_th; a_ph += k_ph; a_pr += k_pr; a_pth += k_pth;
    const R h6 = h * R(1.0 / 6.0);
    r += h6 * a_r;
    th += h6 * a_th;
    ph += h6 * a_ph;
    pr += h6 * a_pr;
    pth += h6 * a_pth;
    if constexpr (T_OUT) *t += h6 * (a_t + k_t);
}

// The reference's damped step (kerr_hamilton.h:165-178): RK4 with h = hs * pole scale.
// K = KerrK (the renderer's RHS) or KerrKDirect (the reference's grouping).
template <class R, bool T_OUT = false, class K = KerrK<R>>
__device__ __forceinline__ void kerr_rk4(R& r, R& th, R& ph, R& pr, R& pth, R hs, const SceneConst<R>& sc,
                                         const K& k, R* t = nullptr) {
    R st, ct;
    dsincos(th, &st, &ct);
    kerr_rk4_raw<R, T_OUT, true, K>(r, th, ph, pr, pth, hs * kerr_pole_scale(st), st, ct, sc, k, t);
}

// Null-constraint re-projection (SING_REFINE_REPROJECT; the reference defines
// the constraint, kerr_hamilton.h:67-94, but never applies it): one Newton
// step on 2H(p_r, p_theta) = 0 along the gradient of 2H in the momenta,
// g = (2 Delta p_r, 2 p_theta) / Sigma.  With 2H = F / Sigma (F the separated
// numerator of kerr_rhs_sc) the step p -= (2H / |g|^2) g reads
//   p_r *= 1 - c Delta,  p_theta *= 1 - c,  c = F / (2 (Delta^2 p_r^2 + p_theta^2)),
// Sigma cancelling; F Delta s2 and the denominator share one reciprocal.
template <class R>
__device__ __forceinline__ void kerr_reproject(R r, R& pr, R& pth, R st, const SceneConst<R>& sc, const KerrK<R>& k) {
    const R a2 = sc.a2;
    const R s2 = st * st;
    const R safe = pos_max(s2, R(1e-10));
    const R r2a2 = r * r + a2;
    const R Del = r * (r - sc.twoM) + a2;
    const R P = r2a2 * k.E - k.aL;
    const R pr2 = pr * pr, pt2 = pth * pth;
    const R G = Del * Del * pr2 + pt2;
    const R Ds = Del * safe;
    const R den = R(2) * G * Ds;
    if (!(den != R(0))) return;
    const R FDs = (((Del * pr2 + pt2) + k.c2aEL) + k.a2E2 * s2) * Ds + (k.L2 * Del - P * P * safe);
    const R c = FDs * drcp(den);
    pr = pr * (R(1) - c * Del);
    pth = pth * (R(1) - c);
}

// Error-controlled step by step doubling: one step of h against two of h/2
// (h = h_cur * pole scale).  The error norm is max(|dr|/(1+|r|), |dtheta|,
// |dphi|)/15.  Rejected (err > tol, fewer than 11 retries, h_cur above
// 1e-3 h_step): h_cur halves and the state is unchanged.  Accepted: the state
// becomes the two-half-step result and h_cur grows by clamp(0.9 (tol/err)^0.2,
// 0.5, 4), capped at 16 h_step.  Returns true when accepted (h_used = the
// step taken).
template <class R> __device__ __forceinline__ R pow_fifth(R x);
template <> __device__ __forceinline__ float pow_fifth<float>(float x) { return exp2f(0.2f * log2f(x)); }
template <> __device__ __forceinline__ double pow_fifth<double>(double x) { return pow(x, 0.2); }

template <class R>
__device__ __forceinline__ bool kerr_adaptive_step(R& r, R& th, R& ph, R& pr, R& pth, R& h_cur, std::uint32_t& tries,
                                                   R tol, const SceneConst<R>& sc, const KerrK<R>& k, R& h_used) {
    R st, ct;
    dsincos(th, &st, &ct);
    const R h = h_cur * kerr_pole_scale(st);
    // FP32: stage sin/cos by rotation; FP64 keeps the full evaluation (the
    // rotation's fallback branches would push the FP64 kernel past 256 VGPRs)
    constexpr bool SH = std::is_same<R, float>::value;
    R br = r, bth = th, bph = ph, bpr = pr, bpt = pth;
    kerr_rk4_raw<R, false, SH>(br, bth, bph, bpr, bpt, h, st, ct, sc, k);
    R hr = r, hth = th, hph = ph, hpr = pr, hpt = pth;
    kerr_rk4_raw<R, false, SH>(hr, hth, hph, hpr, hpt, R(0.5) * h, st, ct, sc, k);
    if constexpr (SH) sincos_shift(th, hth - th, st, ct, &st, &ct);  // the start's, rotated by the exact difference
    else dsincos(hth, &st, &ct);
    kerr_rk4_raw<R, false, SH>(hr, hth, hph, hpr, hpt, R(0.5) * h, st, ct, sc, k);
    const R er = (hr - br) / (R(1) + (hr < R(0) ? -hr : hr));
    const R et = hth - bth, ep = hph - bph;
    const R aer = er < R(0) ? -er : er, aet = et < R(0) ? -et : et, aep = ep < R(0) ? -ep : ep;
    const R err = dmax(dmax(aer, aet), aep) / R(15);
    if (!(err <= tol) && tries < 11u && h_cur > sc.h * R(1e-3)) {
        h_cur = h_cur * R(0.5);
        tries += 1u;
        return false;
    }
    R grow = (err > R(0)) ? R(0.9) * pow_fifth(tol / err) : R(4);
    grow = (grow > R(4)) ? R(4) : ((grow < R(0.5)) ? R(0.5) : grow);
    const R hn = h_cur * grow, cap = sc.h * R(16);
    h_cur = (hn < cap) ? hn : cap;
    tries = 0u;
    h_used = h;
    r = hr;
    th = hth;
    ph = hph;
    pr = hpr;
    pth = hpt;
    return true;
}

// Equator crossing of an error-controlled step: cubic Hermite dense output on
// the step from the end states and their RHS derivatives (step h); three Newton
// iterations on theta(s) = pi/2 from the linear estimate `frac`, then r and phi
// at s.  O(h^4) where the linear interpolation of the fixed-step algorithm is
// O(h^2); error-controlled steps reach 16 h_step far from the hole.
template <class R> __device__ __forceinline__ R hermite(R s, R x0, R d0, R x1, R d1) {
    const R s2 = s * s, s3 = s2 * s;
    return (R(2) * s3 - R(3) * s2 + R(1)) * x0 + (s3 - R(2) * s2 + s) * d0 + (R(3) * s2 - R(2) * s3) * x1 +
           (s3 - s2) * d1;
}
template <class R> __device__ __forceinline__ R hermite_ds(R s, R x0, R d0, R x1, R d1) {
    const R s2 = s * s;
    return (R(6) * s2 - R(6) * s) * (x0 - x1) + (R(3) * s2 - R(4) * s + R(1)) * d0 + (R(3) * s2 - R(2) * s) * d1;
}
// Three Newton iterations on hermite(s) = target from s, clamped to [0, 1].
template <class R> __device__ __forceinline__ R dense_root(R x0, R d0, R x1, R d1, R target, R s) {
    for (int it = 0; it < 3; ++it) {
        const R g = hermite(s, x0, d0, x1, d1) - target;
        const R dg = hermite_ds(s, x0, d0, x1, d1);
        if (dg != R(0)) s = s - g / dg;
        s = (s < R(0)) ? R(0) : ((s > R(1)) ? R(1) : s);
    }
    return s;
}

// The step's end states (r, theta, phi) and their derivatives times h.
template <class R> struct KerrDense {
    R r0, th0, ph0, r1, th1, ph1, dr0, dth0, dph0, dr1, dth1, dph1;
};
template <class R>
__device__ __forceinline__ KerrDense<R> kerr_dense(R r0, R th0, R ph0, R pr0, R pt0, R r1, R th1, R ph1, R pr1,
                                                   R pt1, R h, const SceneConst<R>& sc, const KerrK<R>& k) {
    KerrDense<R> d;
    R st, ct, dpr, dpt;
    dsincos(th0, &st, &ct);
    kerr_rhs_sc(r0, pr0, pt0, st, ct, sc, k, d.dr0, d.dth0, d.dph0, dpr, dpt);
    dsincos(th1, &st, &ct);
    kerr_rhs_sc(r1, pr1, pt1, st, ct, sc, k, d.dr1, d.dth1, d.dph1, dpr, dpt);
    d.r0 = r0; d.th0 = th0; d.ph0 = ph0; d.r1 = r1; d.th1 = th1; d.ph1 = ph1;
    d.dr0 *= h; d.dth0 *= h; d.dph0 *= h; d.dr1 *= h; d.dth1 *= h; d.dph1 *= h;
    return d;
}
template <class R> __device__ __forceinline__ void kerr_dense_crossing(const KerrDense<R>& d, R frac, R& rc, R& phc) {
    const R mid = R(1.5707963267948966192313216916397514);
    const R s = dense_root(d.th0, d.dth0, d.th1, d.dth1, mid, frac);
    rc = hermite(s, d.r0, d.dr0, d.r1, d.dr1);
    phc = hermite(s, d.ph0, d.dph0, d.ph1, d.dph1);
}
// Escape of an error-controlled step: the direction where the dense output
// reaches r = r_esc (the first state beyond it moves by a whole step whenever
// the step sequence changes, e.g. between FP32 and FP64).
template <class R> __device__ __forceinline__ void kerr_dense_escape(const KerrDense<R>& d, R r_esc, R& th, R& ph) {
    const R s = dense_root(d.r0, d.dr0, d.r1, d.dr1, r_esc, (r_esc - d.r0) / (d.r1 - d.r0));
    th = hermite(s, d.th0, d.dth0, d.th1, d.dth1);
    ph = hermite(s, d.ph0, d.dph0, d.ph1, d.dph1);
}

// ---- Schwarzschild geodesic RHS (geodesic_math.h:101-123) ----------------------
// State (r, theta, phi, u^t, u^r, u^theta, u^phi).
template <class R>
__device__ __forceinline__ void schw_rhs(R r, R th, R ut, R ur, R uth, R uph, R rs, R& d_ut, R& d_ur, R& d_uth,
                                         R& d_uph) {
    R st, ct;
    dsincos(th, &st, &ct);
    const R ir = drcp(r);
    const R f = R(1) - rs * ir;
    const R iF = drcp(f);
    const R is = drcp(st);
    const R q = rs * ir * ir;  // rs / r^2
    d_ut = -(q * iF) * ut * ur;
    d_ur = R(-0.5) * (q * f) * ut * ut + R(0.5) * (q * iF) * ur * ur + r * f * (uth * uth + st * st * uph * uph);
    d_uth = R(-2) * ur * uth * ir + st * ct * uph * uph;
    d_uph = R(-2) * ur * uph * ir - R(2) * (ct * is) * uth * uph;
}

// T_OUT: t advances too (*t; d t / d lambda = u^t).
template <class R, bool T_OUT = false>
__device__ __forceinline__ void schw_rk4(R& r, R& th, R& ph, R& ut, R& ur, R& uth, R& uph, R h, R rs,
                                         R* t = nullptr) {
    const R hh = R(0.5) * h;
    R d_ut, d_ur, d_uth, d_uph;
    // stage 1: position derivatives are the velocities themselves
    schw_rhs(r, th, ut, ur, uth, uph, rs, d_ut, d_ur, d_uth, d_uph);
    R a_r = ur, a_th = uth, a_ph = uph, a_ut = d_ut, a_ur = d_ur, a_uth = d_uth, a_uph = d_uph;
    R y_r = r + hh * ur, y_th = th + hh * uth;
    R y_ut = ut + hh * d_ut, y_ur = ur + hh * d_ur, y_uth = uth + hh * d_uth, y_uph = uph + hh * d_uph;
    R a_t = ut;  // d t / d lambda = u^t at each stage (T_OUT only)
    if constexpr (T_OUT) a_t += R(2) * y_ut;
    // stage 2
    schw_rhs(y_r, y_th, y_ut, y_ur, y_uth, y_uph, rs, d_ut, d_ur, d_uth, d_uph);
    a_r += R(2) * y_ur; a_th += R(2) * y_uth; a_ph += R(2) * y_uph;
    a_ut += R(2) * d_ut; a_ur += R(2) * d_ur; a_uth += R(2) * d_uth; a_uph += R(2) * d_uph;
    R n_r = r + hh * y_ur, n_th = th + hh * y_uth;
    R n_ut = ut + hh * d_ut, n_ur = ur + hh * d_ur, n_uth = uth + hh * d_uth, n_uph = uph + hh * d_uph;
    if constexpr (T_OUT) a_t += R(2) * n_ut;
    // stage 3
    schw_rhs(n_r, n_th, n_ut, n_ur, n_uth, n_uph, rs, d_ut, d_ur, d_uth, d_uph);
    a_r += R(2) * n_ur; a_th += R(2) * n_uth; a_ph += R(2) * n_uph;
    a_ut += R(2) * d_ut; a_ur += R(2) * d_ur; a_uth += R(2) * d_uth; a_uph += R(2) * d_uph;
    y_r = r + h * n_ur; y_th = th + h * n_uth;
    y_ut = ut + h * d_ut; y_ur = ur + h * d_ur; y_uth = uth + h * d_uth; y_uph = uph + h * d_uph;
    if constexpr (T_OUT) a_t += y_ut;
    // stage 4
    schw_rhs(y_r, y_th, y_ut, y_ur, y_uth, y_uph, rs, d_ut, d_ur, d_uth, d_uph);
    a_r += y_ur; a_th += y_uth; a_ph += y_uph;
    a_ut += d_ut; a_ur += d_ur; a_uth += d_uth; a_uph += d_uph;
    const R h6 = h * R(1.0 / 6.0);
    if constexpr (T_OUT) *t += h6 * a_t;
    r += h6 * a_r; th += h6 * a_th; ph += h6 * a_ph;
    ut += h6 * a_ut; ur += h6 * a_ur; uth += h6 * a_uth; uph += h6 * a_uph;
}

// ---- shading (geodesic_kernel.cu:77-171) ----------------------------------------
template <class R> __device__ __forceinline__ void blackbody(R T, R* rgb) {
    T = dclamp(T, R(1000), R(40000)) * R(0.01);
    R r, g, b;
    if (T <= R(66)) {
        r = R(1);
        g = dclamp(R(0.39008157876901960784) * dlog(T) - R(0.63184144378862745098), R(0), R(1));
    } else {
        const R t1 = T - R(60);
        r = dclamp(R(1.29293618606274509804) * dpow(t1, R(-0.1332047592)), R(0), R(1));
        g = dclamp(R(1.12989086089529411765) * dpow(t1, R(-0.0755148492)), R(0), R(1));
    }
    if (T >= R(66)) b = R(1);
    else if (T <= R(19)) b = R(0);
    else b = dclamp(R(0.54320678911019607843) * dlog(T - R(10)) - R(1.19625408914), R(0), R(1));
    rgb[0] = r; rgb[1] = g; rgb[2] = b;
}

template <class R> __device__ __forceinline__ R disc_temperature(R rc, R r_in, R peak) {
    if (rc < r_in) return R(0);
    const R ratio = r_in / rc;
    const R q = dpow(ratio, R(0.75));
    const R edge = R(1) - dsqrt(ratio);
    return peak * (q * dmax(edge, R(0))) / R(0.186);
}

template <class R> __device__ __forceinline__ R band_mask(const SceneConst<R>& sc, R rc, R phi) {
    const R x = rc / sc.r_in;
    const R rho = dlog(dmax(x, R(0.1)));
    const R shear = R(4) / dmax(dpow(x, R(1.5)), R(0.25));
    const R omega = dpow(dmax(x, R(0.1)), R(-1.5));
    const R pa = phi - R(0.8) * sc.time_sec * omega;
    const R b1 = R(0.5) + R(0.5) * dcos(R(10) * rho + R(3) * pa + shear + sc.ph1);
    const R b2 = R(0.5) + R(0.5) * dcos(R(22) * rho - R(1.7) * pa + sc.ph2);
    const R combined = R(0.8) + sc.turbulence * (b1 * b1 * b2 - R(0.8));
    return dclamp(R(0.65) + R(0.55) * combined, R(0), R(1.3));
}

__device__ __forceinline__ std::uint32_t mix32(std::uint32_t a, std::uint32_t b) {
    std::uint32_t x = a * 0x9E3779B9u + b;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// Escape colour: starfield in the final direction (geodesic_kernel.cu:135-171,251-259).
template <class R> __device__ __forceinline__ void shade_escape(const SceneConst<R>& sc, R th, R ph, R* rgb) {
    rgb[0] = rgb[1] = rgb[2] = R(0);
    if (!(sc.flags & SING_FLAG_STARFIELD_ON)) return;
    R st, ct, sp, cp;
    dsincos(th, &st, &ct);
    dsincos(ph, &sp, &cp);
    const R x = st * cp, y = st * sp, z = ct;
    const R len = dsqrt(x * x + y * y + z * z);
    if (len < R(1e-6)) {
        rgb[2] = R(0.005);
        return;
    }
    const R il = R(1) / len;
    const R pi = R(3.14159265358979323846);
    const R theta = dacos(dclamp(z * il, R(-1), R(1)));
    const R phi = datan2(y * il, x * il) + pi;
    int ix = int(phi / (R(2) * pi) * R(768)) % 768;
    int iy = int(theta / pi * R(384)) % 384;
    if (ix < 0) ix += 768;
    if (iy < 0) iy += 384;
    const std::uint32_t seed = mix32(std::uint32_t(ix), std::uint32_t(iy));
    R r = R(0.0015), g = R(0.002), b = R(0.005);
    if ((seed & 0xFFu) < 4u) {
        const R br = R((seed >> 8) & 0xFFu) / R(255);
        const R b2 = br * br;
        const R brightness = b2 * b2;
        const R hue = R((seed >> 16) & 0xFFu) / R(255);
        r = dmin(R(1), r + brightness * (R(1) - R(0.5) * hue));
        g = dmin(R(1), g + brightness * (R(0.8) + R(0.2) * hue * (R(1) - hue) * R(4)));
        b = dmin(R(1), b + brightness * (R(0.4) + R(0.6) * hue));
    }
    rgb[0] = r * R(0.6);
    rgb[1] = g * R(0.6);
    rgb[2] = b * R(0.6);
}

// True when the escape direction lies within `eps` cells of a starfield cell
// boundary and a star sits on either side of it: FP32 and FP64 trajectories
// may then land in different cells (geodesic_kernel.cu:135-153 grid).
__device__ __forceinline__ bool starfield_ambiguous(float th, float ph, float eps) {
    float st, ct, sp, cp;
    dsincos(th, &st, &ct);
    dsincos(ph, &sp, &cp);
    const float x = st * cp, y = st * sp, z = ct;
    const float len = sqrtf(x * x + y * y + z * z);
    if (len < 1e-6f) return false;
    const float il = 1.0f / len;
    const float pi = 3.14159265358979323846f;
    const float fx = (atan2f(y * il, x * il) + pi) / (2.0f * pi) * 768.0f;
    const float fy = acosf(fminf(fmaxf(z * il, -1.0f), 1.0f)) / pi * 384.0f;
    const int ix = int(fx) % 768, iy = int(fy) % 384;
    const float dx = fx - floorf(fx), dy = fy - floorf(fy);
    auto star = [](int cx, int cy) {
        cx = (cx % 768 + 768) % 768;
        cy = (cy % 384 + 384) % 384;
        return (mix32(std::uint32_t(cx), std::uint32_t(cy)) & 0xFFu) < 4u;
    };
    const bool here = star(ix, iy);
    if (dx < eps && (here || star(ix - 1, iy))) return true;
    if (dx > 1.0f - eps && (here || star(ix + 1, iy))) return true;
    if (dy < eps && (here || star(ix, iy - 1))) return true;
    if (dy > 1.0f - eps && (here || star(ix, iy + 1))) return true;
    return false;
}

// Disc emission; kerr: cos psi = clamp(L_z/r) and the Kerr Keplerian rate
// (geodesic_kernel.cu:366-394); schwarzschild: cos psi = clamp(-u^phi r)
// (geodesic_kernel.cu:264-289).
template <class R, bool KERR>
__device__ __forceinline__ void shade_disc(const SceneConst<R>& sc, R rc, R phi, R dop_arg, R* rgb) {
    rgb[0] = rgb[1] = rgb[2] = R(0);
    const R T = disc_temperature(rc, sc.r_in, sc.peak_T);
    if (!(T > R(0))) return;
    R f_emit, g_grav = R(1), g_dop = R(1);
    if (KERR) {
        f_emit = R(1) - sc.twoM / rc;
        if (sc.flags & SING_FLAG_REDSHIFT_ON) g_grav = dsqrt(dmax(f_emit, R(1e-6)) / dmax(sc.f_obs, R(1e-6)));
    } else {
        f_emit = R(1) - sc.rs / rc;
        if (sc.flags & SING_FLAG_REDSHIFT_ON) g_grav = dsqrt(dmax(f_emit / sc.f_cam, R(0)));
    }
    if (sc.flags & SING_FLAG_DOPPLER_ON) {
        R om = dsqrt(sc.M / (rc * rc * rc));
        if (KERR) om = om / (R(1) + sc.a * om);
        const R v = rc * om / dsqrt(dmax(f_emit, R(1e-6)));
        const R cos_psi = KERR ? dclamp(dop_arg / rc, R(-1), R(1)) : dclamp(-dop_arg * rc, R(-1), R(1));
        g_dop = dsqrt(dmax(R(1) - v * v, R(0))) / dmax(R(1) - v * cos_psi, R(1e-4));
    }
    const R g = g_grav * g_dop;
    R col[3];
    blackbody(g * T, col);
    const R q = sc.r_in / rc;
    const R flux = dclamp(q * q * q, R(0.05), R(1.5));
    const R bands = band_mask(sc, rc, phi);
    const R g2 = g * g;
    const R w = flux * bands * g2 * g2;
    rgb[0] = col[0] * w;
    rgb[1] = col[1] * w;
    rgb[2] = col[2] * w;
}

// ---- tone map (geodesic_kernel.cu:413-431) ---------------------------------------
template <class R> __device__ __forceinline__ R aces1(R x) {
    return dclamp((x * (R(2.51) * x + R(0.03))) / (x * (R(2.43) * x + R(0.59)) + R(0.14)), R(0), R(1));
}
__device__ __forceinline__ std::uint8_t to_u8(float c) {
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    const float s = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    return std::uint8_t(fminf(fmaxf(s * 255.0f + 0.5f, 0.0f), 255.0f));
}
__device__ __forceinline__ std::uint8_t to_u8(double c) {
    c = fmin(fmax(c, 0.0), 1.0);
    const double s = c <= 0.0031308 ? 12.92 * c : 1.055 * pow(c, 1.0 / 2.4) - 0.055;
    return std::uint8_t(fmin(fmax(s * 255.0 + 0.5, 0.0), 255.0));
}

// ---- camera ray (geodesic_kernel.cu:188-223) -----------------------------------------
template <class R>
__device__ __forceinline__ void camera_ray(const SceneConst<R>& sc, std::uint32_t gx, std::uint32_t gy, R ox, R oy,
                                           R& ur, R& uth, R& uph) {
    const R fx = R(gx) + ox, fy = R(gy) + oy;
    const R un = (R(2) * fx / sc.width_f - R(1)) * sc.aspect * sc.tan_half_fov;
    const R vn = -(R(2) * fy / sc.height_f - R(1)) * sc.tan_half_fov;
    R d0 = sc.fwd[0] + (sc.right[0] * un + sc.up[0] * vn);
    R d1 = sc.fwd[1] + (sc.right[1] * un + sc.up[1] * vn);
    R d2 = sc.fwd[2] + (sc.right[2] * un + sc.up[2] * vn);
    const R len = dsqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const R ik = R(1) / dmax(len, R(1e-20));
    d0 *= ik;
    d1 *= ik;
    d2 *= ik;
    ur = d0 * sc.rhat[0] + d1 * sc.rhat[1] + d2 * sc.rhat[2];
    uth = (d0 * sc.that[0] + d1 * sc.that[1] + d2 * sc.that[2]) / sc.cam_r;
    uph = (d0 * sc.phat[0] + d1 * sc.phat[1] + d2 * sc.phat[2]) / sc.uphi_den;
}

// Photon-orbit margin of a ray with impact parameters xi = L/E, eta = Q/E^2:
// R(r)/E^2 = r^4 + (a^2 - xi^2 - eta) r^2 + 2 M ((xi - a)^2 + eta) r - a^2 eta.
// A ray is near-critical (it winds around the photon shell and amplifies
// rounding exponentially) when the local minimum of R outside the horizon is
// close to zero.  Returns |R(r*)| / r*^4 at the largest stationary point r*,
// or -1 when R has no stationary point outside r+ (no turning point: an
// ingoing ray is captured, an outgoing one escapes directly).
template <class R> __device__ __forceinline__ R photon_margin(R xi, R eta, R a, R M, R r_plus) {
    const R C = (xi - a) * (xi - a) + eta;
    const R P = R(0.5) * (a * a - xi * xi - eta);  // R'/4 = r^3 + P r + Qc
    const R Qc = R(0.5) * M * C;
    if (!(P < R(0))) return R(-1);
    const R disc = -(R(4) * P * P * P + R(27) * Qc * Qc);
    if (!(disc > R(0))) return R(-1);
    const R m = R(2) * dsqrt(-P / R(3));
    const R arg = dclamp((R(3) * Qc / (R(2) * P)) * dsqrt(R(-3) / P), R(-1), R(1));
    const R rs = m * dcos(dacos(arg) / R(3));
    if (!(rs > r_plus)) return R(-1);
    const R r2 = rs * rs;
    const R Rv = r2 * r2 + (a * a - xi * xi - eta) * r2 + R(2) * M * C * rs - a * a * eta;
    const R v = Rv / (r2 * r2);
    return v < R(0) ? -v : v;
}

// Mixed precision, before the march: is this ray hard for FP32?  From its
// constants (xi = L_z/E, eta = Q/E^2) alone: close to a photon orbit (margin
// below band); ingoing and able to reach r = 2 r+ (the radial potential is
// non-negative there: captured rays and rays that turn just outside the hole,
// whose errors the near-horizon metric amplifies); or able to come within
// 0.03 rad of the polar axis (the theta turning point u = cos^2 theta solves
// a^2 u^2 + (eta + xi^2 - a^2) u - eta = 0; near the axis the coordinate
// singularity amplifies rounding and the pole damping makes the ray long).
// Such rays go straight to the FP64 pass, so no FP32 work is wasted on them.
template <class R>
__device__ __forceinline__ bool hard_ray(R xi, R eta, R a, R M, R r_plus, bool ingoing, R band) {
    const R m = photon_margin<R>(xi, eta, a, M, r_plus);
    if (m >= R(0) && m < band) return true;
    const R C = (xi - a) * (xi - a) + eta;
    const R r1 = R(2) * r_plus, r2 = r1 * r1;
    const R Rv = r2 * r2 + (a * a - xi * xi - eta) * r2 + R(2) * M * C * r1 - a * a * eta;
    if (ingoing && Rv >= R(0)) return true;
    if (eta < R(0)) return true;  // vortical orbit: confined to a cone about an axis
    const R B = eta + xi * xi - a * a;
    const R umax = eta > R(0) ? R(2) * eta / (B + dsqrt(B * B + R(4) * a * a * eta)) : R(0);
    return R(1) - umax < R(8.9973e-4);  // sin^2(0.03)
}

// hard_ray's verdict with the kind of ray: 0 not hard; 1 close to a photon
// orbit, vortical or reaching the polar axis (long rays: they wind round the
// hole or are slowed by the pole damping, ~600 steps); 2 only ingoing and able
// to reach 2 r+ (short rays: captured, or meeting the disc near the hole,
// ~70 steps).  The FP64 pass takes the long ones first, so that its last rays
// are short ones and its waves end together.
template <class R>
__device__ __forceinline__ std::uint32_t hard_ray_kind(R xi, R eta, R a, R M, R r_plus, bool ingoing, R band) {
    const R m = photon_margin<R>(xi, eta, a, M, r_plus);
    if (m >= R(0) && m < band) return 1u;
    if (eta < R(0)) return 1u;
    const R B = eta + xi * xi - a * a;
    const R umax = eta > R(0) ? R(2) * eta / (B + dsqrt(B * B + R(4) * a * a * eta)) : R(0);
    if (R(1) - umax < R(8.9973e-4)) return 1u;
    const R C = (xi - a) * (xi - a) + eta;
    const R r1 = R(2) * r_plus, r2 = r1 * r1;
    const R Rv = r2 * r2 + (a * a - xi * xi - eta) * r2 + R(2) * M * C * r1 - a * a * eta;
    return (ingoing && Rv >= R(0)) ? 2u : 0u;
}

}  // namespace sing
