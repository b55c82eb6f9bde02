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
// p_th^2)/Sigma), with one reciprocal.  kerr_rhs_vjp_sc differentiates this
// form.
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

// v^T (d rhs / d(y, E, L, a)) of kerr_rhs_sc, the function the forward
// evaluates: accumulates into gy (no phi dependence) and gc.  (st, ct) =
// sin/cos of y.th as the forward evaluated them.  A reverse sweep of the
// separated form: with 1/Sigma, 1/Delta and 1/s2 as independent quantities
// (d(1/x) = -(1/x)^2 dx), about half the operations of differentiating the
// reference's grouping (B/(Sigma Delta) + ...) of the same function.
template <class T>
SING_HD void kerr_rhs_vjp_sc(const KY<T>& y, const KC<T>& c, T st, T ct, const KY<T>& v, KY<T>& gy, KG<T>& gc) {
    // ---- forward recompute (as kerr_rhs_sc)
    const T a = c.a, a2 = a * a, E = c.E, L = c.L;
    const T aE = a * E;
    const T s2 = st * st;
    const bool clamped = s2 < T(1e-10);
    const T safe = clamped ? T(1e-10) : s2;
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
    const T sc = st * ct;
    // ---- reverse sweep
    // dpt = (sc iS) (L^2 u^2 - (aE)^2 - a2 FiS)
    const T B1 = (L * L * (u * u) - aE * aE) - a2 * FiS;
    const T gB1 = v.pt * (sc * iS);
    T g_sc = v.pt * B1 * iS;
    T g_iS = v.pt * B1 * sc;
    T g_u = gB1 * T(2) * L * L * u;
    T g_FiS = -gB1 * a2;
    T g_L = gB1 * T(2) * L * (u * u);
    T g_aE = -gB1 * T(2) * aE;
    T g_a2 = -gB1 * FiS;
    // dpr = C1 iS, C1 = 2E r PiD - rmM (PiD^2 + pr2) + r FiS
    const T C1 = ((E + E) * r) * PiD - rmM * (PiD * PiD + pr2) + r * FiS;
    const T gC1 = v.pr * iS;
    g_iS += v.pr * C1;
    T g_r = gC1 * (((E + E) * PiD - (PiD * PiD + pr2)) + FiS);
    T g_PiD = gC1 * ((E + E) * r - T(2) * rmM * PiD);
    T g_pr2 = -gC1 * rmM;
    g_FiS += gC1 * r;
    T g_E = gC1 * T(2) * r * PiD;
    // dph = D1 iS, D1 = a PiD + L u - aE
    const T D1 = a * PiD + (L * u - aE);
    const T gD1 = v.ph * iS;
    g_iS += v.ph * D1;
    g_PiD += gD1 * a;
    g_u += gD1 * L;
    g_L += gD1 * u;
    T g_a = gD1 * PiD;
    g_aE -= gD1;
    // dth = pt iS, dr = Del pr iS
    T g_pt = v.th * iS;
    g_iS += v.th * y.pt + v.r * Del * y.pr;
    T g_Del = v.r * y.pr * iS;
    T g_pr = v.r * Del * iS;
    // FiS = F iS
    const T g_F = g_FiS * iS;
    g_iS += g_FiS * F;
    // F = Del pr2 + pt^2 - P PiD + L^2 u - 2 aE L + aE^2 s2
    g_Del += g_F * pr2;
    g_pr2 += g_F * Del;
    g_pt += g_F * T(2) * y.pt;
    T g_P = -g_F * PiD;
    g_PiD -= g_F * P;
    g_u += g_F * L * L;
    g_L += g_F * (T(2) * L * u - T(2) * aE);
    g_aE += g_F * (T(2) * aE * s2 - T(2) * L);
    T g_s2 = g_F * aE * aE;
    // pr2, PiD = P iD, and the reciprocals iS = 1/Sig, iD = 1/Del, u = 1/safe
    g_pr += g_pr2 * T(2) * y.pr;
    g_P += g_PiD * iD;
    const T g_iD = g_PiD * P;
    T g_Sig = -g_iS * iS * iS;
    g_Del -= g_iD * iD * iD;
    const T g_safe = -g_u * u * u;
    // P = r2a2 E - a L
    T g_r2a2 = g_P * E;
    g_E += g_P * r2a2;
    g_a -= g_P * L;
    g_L -= g_P * a;
    // Del = r (r - 2M) + a2, Sig = r2a2 - a2 s2, r2a2 = r^2 + a2
    g_r += g_Del * T(2) * rmM;
    g_a2 += g_Del;
    g_r2a2 += g_Sig;
    g_a2 -= g_Sig * s2;
    g_s2 -= g_Sig * a2;
    g_r += g_r2a2 * T(2) * r;
    g_a2 += g_r2a2;
    // safe = max(s2, 1e-10), s2 = st^2, sc = st ct; aE = a E, a2 = a^2
    if (!clamped) g_s2 += g_safe;
    const T g_st = g_s2 * T(2) * st + g_sc * ct;
    const T g_ct = g_sc * st;
    g_a += g_aE * E + g_a2 * T(2) * a;
    g_E += g_aE * a;
    gc.E += g_E;
    gc.L += g_L;
    gc.a += g_a;
    gy.r += g_r;
    gy.th += g_st * ct - g_ct * st;
    gy.pr += g_pr;
    gy.pt += g_pt;
}
template <class T> SING_HD void kerr_rhs_vjp(const KY<T>& y, const KC<T>& c, const KY<T>& v, KY<T>& gy, KG<T>& gc) {
    T st, ct;
    hsincos(y.th, &st, &ct);
    kerr_rhs_vjp_sc(y, c, st, ct, v, gy, gc);
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

// sin/cos of a stage's theta th0 + d: (s0, c0) rotated by d (hsincos_shift),
// in both precisions.
SING_HD void stage_sincos(float th0, float d, float s0, float c0, float* s, float* c) {
    hsincos_shift(th0, d, s0, c0, s, c);
}
SING_HD void stage_sincos(double th0, double d, double s0, double c0, double* s, double* c) {
    hsincos_shift(th0, d, s0, c0, s, c);
}

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
// The stages are recomputed exactly as kerr_step evaluates them (sin/cos of
// each stage's theta included), and the reverse sweep reuses those sin/cos.
template <class T> SING_HD KY<T> kerr_step_vjp(const KY<T>& y, const KC<T>& c, T hs, const KY<T>& lam, KG<T>& gc) {
    T dh, st1, ct1, st2, ct2, st3, ct3, st4, ct4;
    hsincos(y.th, &st1, &ct1);
    const T h = pole_step_sc(st1, ct1, hs, &dh);
    const T hh = T(0.5) * h;
    const KY<T> k1 = kerr_rhs_sc(y, c, st1, ct1);
    stage_sincos(y.th, hh * k1.th, st1, ct1, &st2, &ct2);
    const KY<T> y2 = ky_axpy(y, k1, hh);
    const KY<T> k2 = kerr_rhs_sc(y2, c, st2, ct2);
    stage_sincos(y.th, hh * k2.th, st1, ct1, &st3, &ct3);
    const KY<T> y3 = ky_axpy(y, k2, hh);
    const KY<T> k3 = kerr_rhs_sc(y3, c, st3, ct3);
    stage_sincos(y.th, h * k3.th, st1, ct1, &st4, &ct4);
    const KY<T> y4 = ky_axpy(y, k3, h);
    const KY<T> k4 = kerr_rhs_sc(y4, c, st4, ct4);
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
    kerr_rhs_vjp_sc(y4, c, st4, ct4, gk4, gy4, gc);
    gy4.ph = T(0);  // f does not depend on phi; y4.ph feeds nothing
    gy = ky_axpy(gy, gy4, T(1));
    gk3 = ky_axpy(gk3, gy4, h);
    gh += ky_dot(gy4, k3);
    // k3 = f(y3), y3 = y + hh k2
    KY<T> gy3 = {T(0), T(0), T(0), T(0), T(0)};
    kerr_rhs_vjp_sc(y3, c, st3, ct3, gk3, gy3, gc);
    gy = ky_axpy(gy, gy3, T(1));
    gk2 = ky_axpy(gk2, gy3, hh);
    gh += T(0.5) * ky_dot(gy3, k2);
    // k2 = f(y2), y2 = y + hh k1
    KY<T> gy2 = {T(0), T(0), T(0), T(0), T(0)};
    kerr_rhs_vjp_sc(y2, c, st2, ct2, gk2, gy2, gc);
    gy = ky_axpy(gy, gy2, T(1));
    gk1 = ky_axpy(gk1, gy2, hh);
    gh += T(0.5) * ky_dot(gy2, k1);
    // k1 = f(y)
    kerr_rhs_vjp_sc(y, c, st1, ct1, gk1, gy, gc);
    // h = pole_step(theta)
    gy.th += gh * dh;
    return gy;
}

}  // namespace sing
