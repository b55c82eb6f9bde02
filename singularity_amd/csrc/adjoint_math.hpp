// singularity_amd/csrc/adjoint_math.hpp — math of the discrete adjoint,
// compiled for both host (unit tests through the diagnostics ABI) and device.
//
//   * Dual<N, T>: forward-mode tangents, used only where a ray evaluates a
//     function once (camera initialisation, disc emission, soft opacity).
//   * kerr_rhs / kerr_rhs_vjp: the Kerr Hamiltonian RHS
//     (shared_shader/kerr_hamilton.h:97-155) and its hand-written reverse-mode
//     vector-Jacobian product, statement by statement in reverse order.
//   * kerr_step / kerr_step_vjp: the damped RK4 step (kerr_hamilton.h:165-178)
//     and its discrete adjoint, including the pole-damping dependence of the
//     step on theta.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define SING_HD __host__ __device__ __forceinline__

namespace sing {

// ---------------------------------------------------------------- Dual numbers
template <int N, class T> struct Dual {
    T v;
    T d[N];
    SING_HD Dual() : v(T(0)) {
        for (int i = 0; i < N; ++i) d[i] = T(0);
    }
    SING_HD Dual(T x) : v(x) {  // NOLINT: constants promote
        for (int i = 0; i < N; ++i) d[i] = T(0);
    }
    SING_HD static Dual var(T x, int k) {
        Dual r(x);
        r.d[k] = T(1);
        return r;
    }
};
template <int N, class T> SING_HD Dual<N, T> dchain(const Dual<N, T>& x, T f, T df) {
    Dual<N, T> r(f);
    for (int i = 0; i < N; ++i) r.d[i] = df * x.d[i];
    return r;
}
template <int N, class T> SING_HD Dual<N, T> operator+(const Dual<N, T>& a, const Dual<N, T>& b) {
    Dual<N, T> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N, class T> SING_HD Dual<N, T> operator-(const Dual<N, T>& a, const Dual<N, T>& b) {
    Dual<N, T> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N, class T> SING_HD Dual<N, T> operator-(const Dual<N, T>& a) {
    Dual<N, T> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N, class T> SING_HD Dual<N, T> operator*(const Dual<N, T>& a, const Dual<N, T>& b) {
    Dual<N, T> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N, class T> SING_HD Dual<N, T> operator/(const Dual<N, T>& a, const Dual<N, T>& b) {
    const T ib = T(1) / b.v;
    Dual<N, T> r(a.v * ib);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N, class T> SING_HD Dual<N, T>& operator+=(Dual<N, T>& a, const Dual<N, T>& b) { return a = a + b; }
template <int N, class T> SING_HD Dual<N, T>& operator*=(Dual<N, T>& a, const Dual<N, T>& b) { return a = a * b; }
template <int N, class T> SING_HD bool operator<(const Dual<N, T>& a, const Dual<N, T>& b) { return a.v < b.v; }
template <int N, class T> SING_HD bool operator>(const Dual<N, T>& a, const Dual<N, T>& b) { return a.v > b.v; }
template <int N, class T> SING_HD bool operator<=(const Dual<N, T>& a, const Dual<N, T>& b) { return a.v <= b.v; }
template <int N, class T> SING_HD bool operator>=(const Dual<N, T>& a, const Dual<N, T>& b) { return a.v >= b.v; }

// scalar helpers shared by T = float / double and Dual
SING_HD float hsqrt(float x) { return sqrtf(x); }
SING_HD double hsqrt(double x) { return sqrt(x); }
SING_HD float hsin(float x) { return sinf(x); }
SING_HD double hsin(double x) { return sin(x); }
SING_HD float hcos(float x) { return cosf(x); }
SING_HD double hcos(double x) { return cos(x); }
SING_HD float hlog(float x) { return logf(x); }
SING_HD double hlog(double x) { return log(x); }
SING_HD float hpow(float x, float y) { return powf(x, y); }
SING_HD double hpow(double x, double y) { return pow(x, y); }
SING_HD float hacos(float x) { return acosf(x); }
SING_HD double hacos(double x) { return acos(x); }
SING_HD float hatan2(float y, float x) { return atan2f(y, x); }
SING_HD double hatan2(double y, double x) { return atan2(y, x); }
SING_HD float htan(float x) { return tanf(x); }
SING_HD double htan(double x) { return tan(x); }
SING_HD float hmax(float a, float b) { return fmaxf(a, b); }
SING_HD double hmax(double a, double b) { return fmax(a, b); }
SING_HD float hmin(float a, float b) { return fminf(a, b); }
SING_HD double hmin(double a, double b) { return fmin(a, b); }
SING_HD float hval(float x) { return x; }
SING_HD double hval(double x) { return x; }
// Reciprocal of the RHS denominators: on the device the hardware estimate
// (v_rcp_f32, 1 ulp) in FP32 and the FP64 estimate refined by two Newton
// steps; the host build (unit tests) divides.
SING_HD float hrcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
SING_HD double hrcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

template <int N, class T> SING_HD Dual<N, T> hsqrt(const Dual<N, T>& x) {
    const T s = hsqrt(x.v);
    return dchain(x, s, s > T(0) ? T(0.5) / s : T(0));
}
template <int N, class T> SING_HD Dual<N, T> hsin(const Dual<N, T>& x) { return dchain(x, hsin(x.v), hcos(x.v)); }
template <int N, class T> SING_HD Dual<N, T> hcos(const Dual<N, T>& x) { return dchain(x, hcos(x.v), -hsin(x.v)); }
template <int N, class T> SING_HD Dual<N, T> hlog(const Dual<N, T>& x) { return dchain(x, hlog(x.v), T(1) / x.v); }
template <int N, class T> SING_HD Dual<N, T> htan(const Dual<N, T>& x) {
    const T t = htan(x.v);
    return dchain(x, t, T(1) + t * t);
}
template <int N, class T> SING_HD Dual<N, T> hpow(const Dual<N, T>& x, const Dual<N, T>& y) {
    const T p = hpow(x.v, y.v);  // constant exponents only on this path
    return dchain(x, p, x.v != T(0) ? y.v * p / x.v : T(0));
}
template <int N, class T> SING_HD Dual<N, T> hacos(const Dual<N, T>& x) {
    const T q = T(1) - x.v * x.v;
    return dchain(x, hacos(x.v), q > T(0) ? T(-1) / hsqrt(q) : T(0));
}
template <int N, class T> SING_HD Dual<N, T> hatan2(const Dual<N, T>& y, const Dual<N, T>& x) {
    Dual<N, T> r(hatan2(y.v, x.v));
    const T iq = T(1) / (x.v * x.v + y.v * y.v);
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * iq;
    return r;
}
template <int N, class T> SING_HD Dual<N, T> hmax(const Dual<N, T>& a, const Dual<N, T>& b) {
    return (a.v >= b.v || b.v != b.v) ? a : b;
}
template <int N, class T> SING_HD Dual<N, T> hmin(const Dual<N, T>& a, const Dual<N, T>& b) {
    return (a.v <= b.v || b.v != b.v) ? a : b;
}
template <int N, class T> SING_HD T hval(const Dual<N, T>& x) { return x.v; }
template <class R> SING_HD R hclamp(R x, R lo, R hi) { return hmin(hmax(x, lo), hi); }

// sin/cos of the polar angle (same reduction and polynomials as the render path)
template <class T> SING_HD void hsincos(T x, T* s, T* c) {
    const T k = rint(x * T(0.636619772367581343075535053490057448));
    T r = fma(-k, T(1.5707963267948966), x);
    r = fma(-k, T(6.123233995736766e-17), r);
    // sums of coefficient x power, highest first (each FMA accumulates into
    // the running sum; see dsincos in physics.hpp)
    const T r2 = r * r, r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4, r10 = r8 * r2, r12 = r8 * r4, r14 = r8 * r6;
    T ps = T(2.8114572543455206e-15) * r14;
    ps = fma(T(-7.647163731819816e-13), r12, ps);
    ps = fma(T(1.6059043836821613e-10), r10, ps);
    ps = fma(T(-2.505210838544172e-08), r8, ps);
    ps = fma(T(2.7557319223985893e-06), r6, ps);
    ps = fma(T(-0.0001984126984126984), r4, ps);
    ps = fma(T(0.008333333333333333), r2, ps);
    ps = ps - T(0.16666666666666666);
    const T sr = fma(r * r2, ps, r);
    T pc = T(4.779477332387385e-14) * r14;
    pc = fma(T(-1.1470745597729725e-11), r12, pc);
    pc = fma(T(2.08767569878681e-09), r10, pc);
    pc = fma(T(-2.755731922398589e-07), r8, pc);
    pc = fma(T(2.48015873015873e-05), r6, pc);
    pc = fma(T(-0.001388888888888889), r4, pc);
    pc = fma(T(0.041666666666666664), r2, pc);
    pc = pc - T(0.5);
    const T cr = fma(r2, pc, T(1));
    const int q = int(k);
    const T ss = (q & 1) ? cr : sr;
    const T cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// FP32: the render path's float reduction (two-term pi/2) and polynomials
// (float truncation), instead of the double-length ones above.
SING_HD void hsincos(float x, float* s, float* c) {
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
    const float ss = (q & 1) ? cr : sr;
    const float cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// sin/cos of th0 + d from (s0, c0) = sin/cos th0 by angle addition, with
// Taylor polynomials of sin d and cos d for |d| <= 0.125 (FP32 to d^5 / d^4,
// FP64 to d^11 / d^10: truncation below 0.1 ulp); a larger |d| evaluates
// sin/cos of th0 + d in full.  Stages 2-4 of a step use it.
SING_HD void hsincos_shift(float th0, float d, float s0, float c0, float* s, float* c) {
    if (fabsf(d) <= 0.125f) {
        const float d2 = d * d;
        const float sd = fmaf(d * d2, fmaf(d2, 8.3333333e-03f, -1.6666667e-01f), d);
        const float cd = fmaf(d2, fmaf(d2, 4.1666668e-02f, -0.5f), 1.0f);
        *s = fmaf(s0, cd, c0 * sd);
        *c = fmaf(c0, cd, -(s0 * sd));
    } else {
        hsincos(th0 + d, s, c);
    }
}
SING_HD void hsincos_shift(double th0, double d, double s0, double c0, double* s, double* c) {
    if (fabs(d) <= 0.125) {
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
        hsincos(th0 + d, s, c);
    }
}

// ------------------------------------------------------------- Kerr RHS + VJP
// State y = (r, theta, phi, p_r, p_theta); constants (E, L_z, a) and M.
template <class T> struct KY {
    T r, th, ph, pr, pt;
};
template <class T> struct KC {
    T E, L, a, M;
};
template <class T> struct KG {  // adjoints of the constants
    T E, L, a;
};

// The RHS in Carter's separated form of the same Hamiltonian, 2H = F/Sigma,
// F = Delta p_r^2 + p_th^2 - P^2/Delta + L^2/s2 - 2 a E L + a^2 E^2 sin^2,
// P = (r^2 + a^2) E - a L (s2 = max(sin^2, 1e-10); r^2 + a^2 - Delta = 2 M r
// makes this the reference's B/(Sigma Delta) + L^2/(Sigma s2) + (Delta p_r^2 +
// p_th^2)/Sigma), with one reciprocal.  kerr_rhs_vjp differentiates the
// reference's grouping of the same function.
template <class T> SING_HD KY<T> kerr_rhs_sc(const KY<T>& y, const KC<T>& c, T st, T ct) {
    const T a = c.a, a2 = a * a, E = c.E, L = c.L;
    const T aE = a * E;
    const T s2 = st * st;
    const T safe = s2 < T(1e-10) ? T(1e-10) : s2;
    const T r = y.r;
    const T r2a2 = r * r + a2;
    const T Sig = r2a2 - a2 * s2;
    const T Del = r * (r - T(2) * c.M) + a2;
    const T P = r2a2 * E - a * L;
    const T SD = Sig * Del;
    const T q = hrcp(SD * safe);
    const T iS = q * (Del * safe), iD = q * (Sig * safe), u = q * SD;
    const T PiD = P * iD;
    const T pr2 = y.pr * y.pr;
    const T F = ((Del * pr2 + y.pt * y.pt) - P * PiD) + (L * L * u - T(2) * aE * L) + aE * aE * s2;
    const T FiS = F * iS;
    const T rmM = r - c.M;
    KY<T> d;
    d.r = Del * y.pr * iS;
    d.th = y.pt * iS;
    d.ph = (a * PiD + (L * u - aE)) * iS;
    d.pr = (((E + E) * r) * PiD - rmM * (PiD * PiD + pr2) + r * FiS) * iS;
    d.pt = ((st * ct) * iS) * ((L * L * (u * u) - aE * aE) - a2 * FiS);
    return d;
}
template <class T> SING_HD KY<T> kerr_rhs(const KY<T>& y, const KC<T>& c) {
    T st, ct;
    hsincos(y.th, &st, &ct);
    return kerr_rhs_sc(y, c, st, ct);
}

// v^T (d rhs / d(y, E, L, a)): accumulates into gy (no phi dependence) and gc.
template <class T> SING_HD void kerr_rhs_vjp(const KY<T>& y, const KC<T>& c, const KY<T>& v, KY<T>& gy, KG<T>& gc) {
    // ---- forward recompute
    T st, ct;
    hsincos(y.th, &st, &ct);
    const T twoM = T(2) * c.M;
    const T a2 = c.a * c.a, EE = c.E * c.E, L2 = c.L * c.L;
    const T cB1 = T(2) * twoM * c.a * c.E * c.L, cB0 = a2 * L2, cphi = twoM * c.a * c.E;
    const T s2 = st * st, c2 = ct * ct;
    const bool clamped = s2 < T(1e-10);
    const T safe = clamped ? T(1e-10) : s2;
    const T r = y.r, r2 = r * r;
    const T Sig = r2 + a2 * c2;
    const T Del = r2 - twoM * r + a2;
    const T r2a2 = r2 + a2;
    const T A = r2a2 * r2a2 - a2 * Del * s2;
    const T iS = hrcp(Sig), iD = hrcp(Del), iss = hrcp(safe);
    const T iSD = iS * iD;
    const T Nph = cphi * r * s2 + (Del - a2 * s2) * c.L;
    const T B = cB1 * r - A * EE - cB0;
    const T twor = T(2) * r;
    const T Del_r = twor - twoM;
    const T A_r = T(2) * twor * r2a2 - a2 * Del_r * s2;
    const T B_r = cB1 - A_r * EE;
    const T sin2t = T(2) * st * ct;
    const T Sig_t = -a2 * sin2t;
    const T B_t = a2 * Del * sin2t * EE;
    const T iS2 = iS * iS;
    const T pr2 = y.pr * y.pr;
    const T Dp2 = Del * pr2 + y.pt * y.pt;
    const T X = twor * Del + Sig * Del_r;
    const T Y = Sig_t * safe + Sig * sin2t;
    const T iSD2 = iSD * iSD, iss2 = iss * iss;

    // ---- reverse sweep
    T g_r = T(0), g_pr = T(0), g_pt = T(0), g_st = T(0), g_ct = T(0);
    T g_a2 = T(0), g_EE = T(0), g_L2 = T(0), g_cB1 = T(0), g_cB0 = T(0), g_cphi = T(0), g_L = T(0);
    T g_s2 = T(0), g_c2 = T(0), g_safe = T(0), g_r2 = T(0), g_Sig = T(0), g_Del = T(0), g_r2a2 = T(0), g_A = T(0);
    T g_iS = T(0), g_iD = T(0), g_iss = T(0), g_iSD = T(0), g_B = T(0), g_twor = T(0), g_Del_r = T(0);
    T g_A_r = T(0), g_B_r = T(0), g_sin2t = T(0), g_Sig_t = T(0), g_B_t = T(0), g_iS2 = T(0), g_pr2 = T(0);
    T g_Dp2 = T(0), g_X = T(0), g_Y = T(0), g_Nph = T(0);

    // d.pr = -0.5 dHdr ; d.pt = -0.5 dHdth
    const T gHr = T(-0.5) * v.pr;
    const T gHt = T(-0.5) * v.pt;
    // dHdth = B_t iSD - B Sig_t iS2 iD - L2 Y iS2 iss^2 - Dp2 Sig_t iS2
    g_B_t += gHt * iSD;
    g_iSD += gHt * B_t;
    g_B -= gHt * Sig_t * iS2 * iD;
    g_Sig_t -= gHt * B * iS2 * iD;
    g_iS2 -= gHt * B * Sig_t * iD;
    g_iD -= gHt * B * Sig_t * iS2;
    g_L2 -= gHt * Y * iS2 * iss2;
    g_Y -= gHt * L2 * iS2 * iss2;
    g_iS2 -= gHt * L2 * Y * iss2;
    g_iss -= gHt * L2 * Y * iS2 * T(2) * iss;
    g_Dp2 -= gHt * Sig_t * iS2;
    g_Sig_t -= gHt * Dp2 * iS2;
    g_iS2 -= gHt * Dp2 * Sig_t;
    // dHdr = B_r iSD - B X iSD^2 - L2 twor iS2 iss + Del_r pr2 iS - Dp2 twor iS2
    g_B_r += gHr * iSD;
    g_iSD += gHr * B_r;
    g_B -= gHr * X * iSD2;
    g_X -= gHr * B * iSD2;
    g_iSD -= gHr * B * X * T(2) * iSD;
    g_L2 -= gHr * twor * iS2 * iss;
    g_twor -= gHr * L2 * iS2 * iss;
    g_iS2 -= gHr * L2 * twor * iss;
    g_iss -= gHr * L2 * twor * iS2;
    g_Del_r += gHr * pr2 * iS;
    g_pr2 += gHr * Del_r * iS;
    g_iS += gHr * Del_r * pr2;
    g_Dp2 -= gHr * twor * iS2;
    g_twor -= gHr * Dp2 * iS2;
    g_iS2 -= gHr * Dp2 * twor;
    // Y = Sig_t safe + Sig sin2t
    g_Sig_t += g_Y * safe;
    g_safe += g_Y * Sig_t;
    g_Sig += g_Y * sin2t;
    g_sin2t += g_Y * Sig;
    // X = twor Del + Sig Del_r
    g_twor += g_X * Del;
    g_Del += g_X * twor;
    g_Sig += g_X * Del_r;
    g_Del_r += g_X * Sig;
    // Dp2 = Del pr2 + pt^2
    g_Del += g_Dp2 * pr2;
    g_pr2 += g_Dp2 * Del;
    g_pt += g_Dp2 * T(2) * y.pt;
    // pr2 = pr^2
    g_pr += g_pr2 * T(2) * y.pr;
    // iS2 = iS^2
    g_iS += g_iS2 * T(2) * iS;
    // B_t = a2 Del sin2t EE
    g_a2 += g_B_t * Del * sin2t * EE;
    g_Del += g_B_t * a2 * sin2t * EE;
    g_sin2t += g_B_t * a2 * Del * EE;
    g_EE += g_B_t * a2 * Del * sin2t;
    // Sig_t = -a2 sin2t
    g_a2 -= g_Sig_t * sin2t;
    g_sin2t -= g_Sig_t * a2;
    // sin2t = 2 st ct
    g_st += g_sin2t * T(2) * ct;
    g_ct += g_sin2t * T(2) * st;
    // B_r = cB1 - A_r EE
    g_cB1 += g_B_r;
    g_A_r -= g_B_r * EE;
    g_EE -= g_B_r * A_r;
    // A_r = 2 twor r2a2 - a2 Del_r s2
    g_twor += g_A_r * T(2) * r2a2;
    g_r2a2 += g_A_r * T(2) * twor;
    g_a2 -= g_A_r * Del_r * s2;
    g_Del_r -= g_A_r * a2 * s2;
    g_s2 -= g_A_r * a2 * Del_r;
    // Del_r = twor - twoM
    g_twor += g_Del_r;
    // twor = 2 r
    g_r += g_twor * T(2);
    // B = cB1 r - A EE - cB0
    g_cB1 += g_B * r;
    g_r += g_B * cB1;
    g_A -= g_B * EE;
    g_EE -= g_B * A;
    g_cB0 -= g_B;
    // d.ph = Nph iSD iss
    g_Nph += v.ph * iSD * iss;
    g_iSD += v.ph * Nph * iss;
    g_iss += v.ph * Nph * iSD;
    // Nph = cphi r s2 + (Del - a2 s2) L
    g_cphi += g_Nph * r * s2;
    g_r += g_Nph * cphi * s2;
    g_s2 += g_Nph * (cphi * r - a2 * c.L);
    g_Del += g_Nph * c.L;
    g_a2 -= g_Nph * s2 * c.L;
    g_L += g_Nph * (Del - a2 * s2);
    // d.th = pt iS
    g_pt += v.th * iS;
    g_iS += v.th * y.pt;
    // d.r = Del pr iS
    g_Del += v.r * y.pr * iS;
    g_pr += v.r * Del * iS;
    g_iS += v.r * Del * y.pr;
    // iSD = iS iD
    g_iS += g_iSD * iD;
    g_iD += g_iSD * iS;
    // reciprocals
    g_Sig -= g_iS * iS * iS;
    g_Del -= g_iD * iD * iD;
    g_safe -= g_iss * iss * iss;
    // A = r2a2^2 - a2 Del s2
    g_r2a2 += g_A * T(2) * r2a2;
    g_a2 -= g_A * Del * s2;
    g_Del -= g_A * a2 * s2;
    g_s2 -= g_A * a2 * Del;
    // r2a2 = r2 + a2
    g_r2 += g_r2a2;
    g_a2 += g_r2a2;
    // Del = r2 - twoM r + a2
    g_r2 += g_Del;
    g_r -= g_Del * twoM;
    g_a2 += g_Del;
    // Sig = r2 + a2 c2
    g_r2 += g_Sig;
    g_a2 += g_Sig * c2;
    g_c2 += g_Sig * a2;
    // r2 = r^2
    g_r += g_r2 * T(2) * r;
    // safe = max(s2, 1e-10)
    if (!clamped) g_s2 += g_safe;
    // s2 = st^2, c2 = ct^2
    g_st += g_s2 * T(2) * st;
    g_ct += g_c2 * T(2) * ct;
    // constants: a2 = a^2, EE = E^2, L2 = L^2, cB1 = 2 twoM a E L, cB0 = a2 L2, cphi = twoM a E
    g_a2 += g_cB0 * L2;
    g_L2 += g_cB0 * a2;
    g_L += g_L2 * T(2) * c.L + g_cB1 * T(2) * twoM * c.a * c.E;
    gc.L += g_L;
    gc.E += g_EE * T(2) * c.E + g_cB1 * T(2) * twoM * c.a * c.L + g_cphi * twoM * c.a;
    gc.a += g_a2 * T(2) * c.a + g_cB1 * T(2) * twoM * c.E * c.L + g_cphi * twoM * c.E;
    // theta through (st, ct)
    gy.r += g_r;
    gy.th += g_st * ct - g_ct * st;
    gy.pr += g_pr;
    gy.pt += g_pt;
}

template <class T> SING_HD KY<T> ky_axpy(const KY<T>& y, const KY<T>& k, T h) {
    return {y.r + h * k.r, y.th + h * k.th, y.ph + h * k.ph, y.pr + h * k.pr, y.pt + h * k.pt};
}
template <class T> SING_HD T ky_dot(const KY<T>& a, const KY<T>& b) {
    return a.r * b.r + a.th * b.th + a.ph * b.ph + a.pr * b.pr + a.pt * b.pt;
}

// Pole damping h = hs * clamp(5 |sin theta|, 0.05, 1) (kerr_hamilton.h:166-170) and d h / d theta.
template <class T> SING_HD T pole_step_sc(T st, T ct, T hs, T* dh_dth) {
    const T as = st < T(0) ? -st : st;
    const T ps = as * T(5);
    if (ps > T(1)) {
        *dh_dth = T(0);
        return hs;
    }
    if (ps < T(0.05)) {
        *dh_dth = T(0);
        return hs * T(0.05);
    }
    *dh_dth = hs * T(5) * (st < T(0) ? -ct : ct);
    return hs * ps;
}
template <class T> SING_HD T pole_step(T th, T hs, T* dh_dth) {
    T st, ct;
    hsincos(th, &st, &ct);
    return pole_step_sc(st, ct, hs, dh_dth);
}

// sin/cos of a stage's theta th0 + d: FP32 rotates (s0, c0) (hsincos_shift);
// FP64 evaluates in full (the rotation's registers push the FP64 backward
// kernel, which recomputes steps inside the VJP sweep, into more spills).
SING_HD void stage_sincos(float th0, float d, float s0, float c0, float* s, float* c) {
    hsincos_shift(th0, d, s0, c0, s, c);
}
SING_HD void stage_sincos(double th0, double d, double, double, double* s, double* c) { hsincos(th0 + d, s, c); }

// The damped RK4 step.  sin/cos of theta are evaluated once, shared by the
// pole damping and the first stage.
template <class T> SING_HD KY<T> kerr_step(const KY<T>& y, const KC<T>& c, T hs) {
    T dh, st, ct, sn, cs;
    hsincos(y.th, &st, &ct);
    const T h = pole_step_sc(st, ct, hs, &dh);
    const T hh = T(0.5) * h;
    const KY<T> k1 = kerr_rhs_sc(y, c, st, ct);
    stage_sincos(y.th, hh * k1.th, st, ct, &sn, &cs);
    const KY<T> k2 = kerr_rhs_sc(ky_axpy(y, k1, hh), c, sn, cs);
    stage_sincos(y.th, hh * k2.th, st, ct, &sn, &cs);
    const KY<T> k3 = kerr_rhs_sc(ky_axpy(y, k2, hh), c, sn, cs);
    stage_sincos(y.th, h * k3.th, st, ct, &sn, &cs);
    const KY<T> k4 = kerr_rhs_sc(ky_axpy(y, k3, h), c, sn, cs);
    KY<T> s;
    s.r = k1.r + T(2) * k2.r + T(2) * k3.r + k4.r;
    s.th = k1.th + T(2) * k2.th + T(2) * k3.th + k4.th;
    s.ph = k1.ph + T(2) * k2.ph + T(2) * k3.ph + k4.ph;
    s.pr = k1.pr + T(2) * k2.pr + T(2) * k3.pr + k4.pr;
    s.pt = k1.pt + T(2) * k2.pt + T(2) * k3.pt + k4.pt;
    return ky_axpy(y, s, h * T(1.0 / 6.0));
}

// Discrete adjoint of one step: given lam = dJ/d y_{n+1}, returns dJ/d y_n and
// accumulates dJ/d(E, L, a) into gc.
template <class T> SING_HD KY<T> kerr_step_vjp(const KY<T>& y, const KC<T>& c, T hs, const KY<T>& lam, KG<T>& gc) {
    T dh;
    const T h = pole_step(y.th, hs, &dh);
    const T hh = T(0.5) * h;
    const KY<T> k1 = kerr_rhs(y, c);
    const KY<T> y2 = ky_axpy(y, k1, hh);
    const KY<T> k2 = kerr_rhs(y2, c);
    const KY<T> y3 = ky_axpy(y, k2, hh);
    const KY<T> k3 = kerr_rhs(y3, c);
    const KY<T> y4 = ky_axpy(y, k3, h);
    const KY<T> k4 = kerr_rhs(y4, c);
    // y' = y + h/6 (k1 + 2 k2 + 2 k3 + k4)
    KY<T> gy = lam;
    const T h6 = h * T(1.0 / 6.0), h3 = h * T(1.0 / 3.0);
    KY<T> gk1 = {h6 * lam.r, h6 * lam.th, h6 * lam.ph, h6 * lam.pr, h6 * lam.pt};
    KY<T> gk2 = {h3 * lam.r, h3 * lam.th, h3 * lam.ph, h3 * lam.pr, h3 * lam.pt};
    KY<T> gk3 = gk2;
    const KY<T> gk4 = gk1;
    KY<T> sum;
    sum.r = k1.r + T(2) * k2.r + T(2) * k3.r + k4.r;
    sum.th = k1.th + T(2) * k2.th + T(2) * k3.th + k4.th;
    sum.ph = k1.ph + T(2) * k2.ph + T(2) * k3.ph + k4.ph;
    sum.pr = k1.pr + T(2) * k2.pr + T(2) * k3.pr + k4.pr;
    sum.pt = k1.pt + T(2) * k2.pt + T(2) * k3.pt + k4.pt;
    T gh = ky_dot(lam, sum) * T(1.0 / 6.0);
    // k4 = f(y4), y4 = y + h k3
    KY<T> gy4 = {T(0), T(0), T(0), T(0), T(0)};
    kerr_rhs_vjp(y4, c, gk4, gy4, gc);
    gy4.ph = T(0);  // f does not depend on phi; y4.ph feeds nothing
    gy = ky_axpy(gy, gy4, T(1));
    gk3 = ky_axpy(gk3, gy4, h);
    gh += ky_dot(gy4, k3);
    // k3 = f(y3), y3 = y + hh k2
    KY<T> gy3 = {T(0), T(0), T(0), T(0), T(0)};
    kerr_rhs_vjp(y3, c, gk3, gy3, gc);
    gy = ky_axpy(gy, gy3, T(1));
    gk2 = ky_axpy(gk2, gy3, hh);
    gh += T(0.5) * ky_dot(gy3, k2);
    // k2 = f(y2), y2 = y + hh k1
    KY<T> gy2 = {T(0), T(0), T(0), T(0), T(0)};
    kerr_rhs_vjp(y2, c, gk2, gy2, gc);
    gy = ky_axpy(gy, gy2, T(1));
    gk1 = ky_axpy(gk1, gy2, hh);
    gh += T(0.5) * ky_dot(gy2, k1);
    // k1 = f(y)
    kerr_rhs_vjp(y, c, gk1, gy, gc);
    // h = pole_step(theta)
    gy.th += gh * dh;
    return gy;
}

}  // namespace sing
