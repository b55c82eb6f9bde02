"""Pins of the oracle against mathematics the reference's tests fix.

Each test is chosen so a plausible transcription mistake (a dropped term, a
wrong sign, a transposed operand) fails it:

* Kerr RHS (FP64) equals SymPy's own derivative of 2H (the form of
  verification/test_kerr_ham_analytic.py:29-48) on its grid, rel 1e-9.
* Schwarzschild RHS equals -Gamma^mu_{nu sigma} u^nu u^sigma with the
  Christoffels derived here by SymPy from the metric
  (verification/test_christoffel_sympy.py).
* RK4 is 4th order (error ratio ~16 when h halves).
* Circular photon orbits close (tests/test_kerr_hamilton.cpp:86-118).
* Ray deflection matches the exact integral (and 4M/b for large b).
* Conservation drifts of E, L, null residual (tests/test_conserved.cpp:45-94).
* Kerr horizon / ISCO tables (tests/test_kerr.cpp:34-39,123-128).
* Crossing known answers (tests/test_disc_intersection.cpp:28-95).
* Halton, disc temperature peak, blackbody, ACES, sRGB and hash known answers.
"""

from __future__ import annotations

import ctypes as C
import itertools
import math

import numpy as np
import pytest
import sympy as sp

from helpers import config1, scene_uniforms

import singularity_amd as sa


def p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- Kerr RHS
_r, _th, _pr, _pt, _E, _L, _a, _M = sp.symbols("r theta p_r p_theta E L a M", real=True)
_s2 = sp.sin(_th) ** 2
_Sig = _r**2 + _a**2 * sp.cos(_th) ** 2
_Del = _r**2 - 2 * _M * _r + _a**2
_g_inv_tt = -((_r**2 + _a**2) ** 2 - _a**2 * _Del * _s2) / (_Sig * _Del)
_g_inv_tp = -2 * _M * _r * _a / (_Sig * _Del)
_g_inv_pp = (_Del - _a**2 * _s2) / (_Sig * _Del * _s2)
# 2H = g^{mu nu} p_mu p_nu with p_t = -E, p_phi = L (Boyer-Lindquist inverse metric)
_twoH = (_g_inv_tt * _E**2 - 2 * _g_inv_tp * _E * _L + _g_inv_pp * _L**2 + _Del / _Sig * _pr**2 + _pt**2 / _Sig)
_args = (_r, _th, _pr, _pt, _E, _L, _a, _M)
_d_r = sp.lambdify(_args, sp.diff(_twoH, _r), "numpy")
_d_th = sp.lambdify(_args, sp.diff(_twoH, _th), "numpy")
_d_pr = sp.lambdify(_args, sp.diff(_twoH, _pr), "numpy")
_d_pt = sp.lambdify(_args, sp.diff(_twoH, _pt), "numpy")
_d_L = sp.lambdify(_args, sp.diff(_twoH, _L), "numpy")
_d_E = sp.lambdify(_args, sp.diff(_twoH, _E), "numpy")

GRID = list(itertools.product([2.6, 3.1, 6.0, 12.0, 40.0, 150.0],
                              [0.05, 0.4, 1.0, np.pi / 2, 2.4, np.pi - 0.05],
                              [0.0, 0.5, 0.94, 0.998], [-7.5, 0.0, 3.7]))


def test_kerr_rhs_is_hamiltons_equations_sympy(orc):
    L = orc.strict()
    d = np.zeros(6)
    for r, th, a, Lz in GRID:
        E, M, pr, pt = 1.0, 1.0, 0.6, 2.3
        s = np.array([0.0, r, th, 0.3, pr, pt])
        c = np.array([E, Lz, 0.0, a, M])
        L.orc_kerr_rhs_f64(p(s), p(c), p(d))
        args = (r, th, pr, pt, E, Lz, a, M)
        # dx/dlam = dH/dp = 1/2 d(2H)/dp ; dp/dlam = -1/2 d(2H)/dx
        want = {
            "r": 0.5 * _d_pr(*args), "theta": 0.5 * _d_pt(*args), "phi": 0.5 * _d_L(*args),
            "t": -0.5 * _d_E(*args), "p_r": -0.5 * _d_r(*args), "p_theta": -0.5 * _d_th(*args),
        }
        got = {"t": d[0], "r": d[1], "theta": d[2], "phi": d[3], "p_r": d[4], "p_theta": d[5]}
        for k in want:
            assert got[k] == pytest.approx(float(want[k]), rel=1e-9, abs=1e-9), (k, r, th, a, Lz)


def test_kerr_two_H_zero_on_null_camera_rays(orc):
    """2H = 0 for the rays the camera launches (oracle trace invariants in FP64)."""
    u = scene_uniforms(spin=0.94, r_in=1.7, r_out=30.0, width=16, height=16)
    for gx, gy in [(0, 0), (8, 8), (15, 3), (4, 12)]:
        r = orc.trace_ray(u, gx, gy, precision=64)
        assert abs(r["inv0"][0]) < 1e-10


# ------------------------------------------------------- Schwarzschild RHS
def test_schwarzschild_rhs_matches_sympy_christoffels(orc):
    t, r, th, ph = sp.symbols("t r theta phi", real=True)
    rs = sp.Symbol("r_s", positive=True)
    x = [t, r, th, ph]
    f = 1 - rs / r
    g = sp.diag(-f, 1 / f, r**2, r**2 * sp.sin(th) ** 2)
    gi = g.inv()
    Gam = [[[sp.simplify(sum(gi[m, k] * (sp.diff(g[k, n], x[s]) + sp.diff(g[k, s], x[n]) - sp.diff(g[n, s], x[k]))
                                 for k in range(4)) / 2) for s in range(4)] for n in range(4)] for m in range(4)]
    u = sp.symbols("u0:4", real=True)
    acc = [sp.lambdify((r, th, rs, *u), -sum(Gam[m][n][s] * u[n] * u[s] for n in range(4) for s in range(4)), "numpy")
           for m in range(4)]
    L = orc.strict()
    rng = np.random.default_rng(3)
    d = np.zeros(8)
    for _ in range(200):
        st = np.array([0.0, rng.uniform(2.2, 60), rng.uniform(0.1, 3.0), rng.uniform(0, 6), rng.uniform(1, 2),
                       rng.normal(0, 1), rng.normal(0, 0.05), rng.normal(0, 0.05)])
        L.orc_schw_rhs_f64(p(st), 2.0, p(d))
        assert d[0] == st[4] and d[1] == st[5] and d[2] == st[6] and d[3] == st[7]
        for m in range(4):
            want = float(acc[m](st[1], st[2], 2.0, *st[4:8]))
            assert d[4 + m] == pytest.approx(want, rel=1e-10, abs=1e-12)


# ------------------------------------------------------------ integrators
def test_kerr_rk4_is_fourth_order(orc):
    L = orc.strict()
    s0 = np.array([0.0, 12.0, 1.2, 0.0, -0.4, 1.5])
    c = np.array([1.0, 2.0, 0.0, 0.7, 1.0])
    span = 6.4

    def end(h):
        n = int(round(span / h))
        tr = np.zeros((n + 1, 6))
        L.orc_kerr_rk4_f64(p(s0), h, p(c), n, p(tr))
        return tr[-1]

    ref = end(0.0125)
    e1 = np.abs(end(0.2) - ref).max()
    e2 = np.abs(end(0.1) - ref).max()
    assert 12.0 < e1 / e2 < 20.0


def test_schwarzschild_photon_sphere_closes(orc):
    """Circular photon orbit at r=3M stays within 0.5 % over one period (float)."""
    L = orc.strict()
    r0 = 3.0
    f = 1 - 2.0 / r0
    s = np.array([0, r0, np.pi / 2, 0, 1.0, 0, 0, np.sqrt(f) / r0], np.float32)
    period = 2 * np.pi / s[7]
    h = 0.01
    n = int(period / h)
    tr = np.zeros((n + 1, 8), np.float32)
    L.orc_schw_rk4_f32(p(s), h, 2.0, n, p(tr))
    assert abs(tr[-1, 1] - r0) < 0.005 * r0
    assert abs(tr[-1, 2] - np.pi / 2) < 1e-4


def test_kerr_prograde_photon_orbit_closes(orc):
    """Equatorial prograde circular photon orbit r_ph with b = (r sqrt(Delta) - 2 M a)/(r - 2M)
    (cli/main.cpp:1359); FP64 trajectory keeps r within 0.5 % for one revolution."""
    L = orc.strict()
    a, M = 0.5, 1.0
    r = orc.kerr_radius(4, M, a)
    Del = r * r - 2 * M * r + a * a
    b = (r * np.sqrt(Del) - 2 * M * a) / (r - 2 * M)
    s = np.array([0.0, r, np.pi / 2, 0.0, 0.0, 0.0])
    c = np.array([1.0, b, 0.0, a, M])
    d = np.zeros(6)
    L.orc_kerr_rhs_f64(p(s), p(c), p(d))
    assert abs(d[4]) < 1e-9  # dp_r/dlam = 0 on the photon orbit
    n = int(2 * np.pi / d[3] / 0.01)
    tr = np.zeros((n + 1, 6))
    L.orc_kerr_rk4_f64(p(s), 0.01, p(c), n, p(tr))
    assert np.abs(tr[:, 1] - r).max() < 0.005 * r


def test_deflection_matches_exact_integral(orc):
    """Swept azimuth of an equatorial ray between two passes of r0 equals the exact
    integral 2 int_{u0}^{u_t} du / sqrt(1/b^2 - u^2 + 2 M u^3) (u = 1/r), computed here
    with scipy (the form of verification/test_phase1_rays.py:81-110); for large b
    it reduces to the Einstein angle 4M/b (verification/test_deflection.py)."""
    from scipy.integrate import quad

    L = orc.strict()
    for b in (8.0, 20.0, 120.0):
        r0 = 400.0
        u0 = 1.0 / r0
        F = lambda u: 1.0 / b**2 - u * u + 2.0 * u**3  # noqa: E731
        roots = np.roots([2.0, -1.0, 0.0, 1.0 / b**2])
        ut = min(x.real for x in roots if abs(x.imag) < 1e-12 and x.real > 0)
        dF = -2 * ut + 6 * ut * ut
        g = lambda t: 2 * t / np.sqrt(max(F(ut - t * t), 1e-300)) if t > 0 else 2 / np.sqrt(-dF)  # noqa: E731
        exact = 2 * quad(g, 0.0, np.sqrt(ut - u0), limit=400, epsabs=1e-13, epsrel=1e-12)[0]
        f = 1 - 2.0 / r0
        ur = -np.sqrt(max(1.0 - f * b * b / r0**2, 0.0))  # E = 1, u^t = 1/f, u^phi = b/r^2
        s = np.array([0, r0, np.pi / 2, 0.0, 1 / f, ur, 0, b / r0**2])
        n = 20000
        tr = np.zeros((n + 1, 8))
        L.orc_schw_rk4_f64(p(s), 0.05, 2.0, n, p(tr))
        turned = np.argmin(tr[:, 1])
        k = turned + int(np.argmax(tr[turned:, 1] > r0))
        w = (r0 - tr[k - 1, 1]) / (tr[k, 1] - tr[k - 1, 1])
        swept = tr[k - 1, 3] + w * (tr[k, 3] - tr[k - 1, 3])
        assert swept == pytest.approx(exact, rel=1e-6), b
        if b > 100:
            defl = swept - (np.pi - 2 * np.arcsin(b / r0))
            assert defl == pytest.approx(4.0 / b, rel=0.05)


def test_conservation_drifts_float_and_double(orc):
    """E, L and the null residual drift along C1 rays; caps from tests/test_conserved.cpp:92-94."""
    u = config1(64, 64)
    for gx, gy in [(10, 10), (32, 40), (50, 20), (5, 60)]:
        r32 = orc.trace_ray(u, gx, gy, precision=32)
        r64 = orc.trace_ray(u, gx, gy, precision=64)
        for r, cap in ((r32, 3e-4), (r64, 1e-9)):
            e0, l0, n0 = r["inv0"]
            e1, l1, n1 = r["inv1"]
            if r["cls"] == 0:  # captured rays end inside the coordinate singularity region
                continue
            assert abs(e1 - e0) / abs(e0) < cap
            assert abs(l1 - l0) / max(abs(l0), 1e-6) < cap
            assert abs(n1 - n0) < cap * 10


def test_kerr_carter_constant_drift_double(orc):
    u = scene_uniforms(spin=0.94, r_in=1.7, r_out=30.0, width=32, height=32)
    for gx, gy in [(3, 3), (16, 2), (28, 30), (30, 16)]:
        r = orc.trace_ray(u, gx, gy, precision=64)
        if r["cls"] == 1:
            assert abs(r["inv1"][1] - r["inv0"][1]) < 1e-6 * max(1.0, abs(r["inv0"][1]))
            assert abs(r["inv1"][0]) < 1e-6


# ---------------------------------------------------------- closed forms
@pytest.mark.parametrize("a,rp,rm,isco_p,isco_r", [
    (0.0, 2.0, 0.0, 6.0, 6.0),
    (0.5, 1.866025, 0.133975, 4.233003, 7.554585),
    (0.94, 1.341204, 0.658796, 2.023593, 8.830752),
    (0.998, 1.063246, 0.936754, 1.236971, 8.994374),
])
def test_kerr_tables(orc, a, rp, rm, isco_p, isco_r):
    # the reference's table carries ~3e-5 slack on r+/r- (its tolerance is rel 1e-4);
    # Delta(r+-) = 0 pins the horizons exactly
    assert orc.kerr_radius(0, 1.0, a) == pytest.approx(rp, rel=1e-4)
    assert orc.kerr_radius(1, 1.0, a) == pytest.approx(rm, rel=1e-4, abs=1e-6)
    for w in (0, 1):
        rh = orc.kerr_radius(w, 1.0, a)
        assert abs(rh * rh - 2 * rh + a * a) < 1e-12
    assert orc.kerr_radius(2, 1.0, a) == pytest.approx(isco_p, rel=1e-6)
    assert orc.kerr_radius(3, 1.0, a) == pytest.approx(isco_r, rel=1e-6)
    # photon orbits: 3M at a=0; prograde -> M and retrograde -> 4M as a -> M
    if a == 0.0:
        assert orc.kerr_radius(4, 1.0, a) == pytest.approx(3.0) and orc.kerr_radius(5, 1.0, a) == pytest.approx(3.0)
    # the product's closed forms agree
    assert sa.kerr_radius(2, 1.0, a) == pytest.approx(orc.kerr_radius(2, 1.0, a), rel=1e-12)
    assert sa.kerr_radius(4, 1.0, a) == pytest.approx(orc.kerr_radius(4, 1.0, a), rel=1e-12)


def test_crossing_known_answers(orc):
    L = orc.strict()
    hp = np.float32(np.pi / 2)

    def cross(a, b):
        c, f = C.c_int(0), C.c_float(0)
        L.orc_crossing_f32(float(a), float(b), C.byref(c), C.byref(f))
        return c.value, f.value

    assert cross(hp - np.float32(0.3), hp - np.float32(0.05))[0] == 0
    assert cross(hp + np.float32(0.2), hp + np.float32(0.4))[0] == 0
    c, f = cross(hp - np.float32(0.1), hp + np.float32(0.1))
    assert c == 1 and f == pytest.approx(0.5, abs=1e-6)
    c, f = cross(hp - np.float32(0.04), hp + np.float32(0.12))
    assert c == 1 and f == pytest.approx(0.25, abs=1e-6)
    c1, f1 = cross(hp - np.float32(0.2), hp + np.float32(0.3))
    c2, f2 = cross(hp + np.float32(0.3), hp - np.float32(0.2))
    assert c1 and c2 and f1 + f2 == pytest.approx(1.0, abs=1e-6)
    c, f = cross(hp - np.float32(0.15), hp + np.float32(0.05))
    assert 12.4 + f * (11.9 - 12.4) == pytest.approx(12.025, abs=1e-5)


def test_halton_known_answers(orc):
    L = orc.strict()
    assert L.orc_halton(1, 2) == 0.5 and L.orc_halton(2, 2) == 0.25 and L.orc_halton(3, 2) == 0.75
    assert L.orc_halton(1, 3) == pytest.approx(1 / 3) and L.orc_halton(2, 3) == pytest.approx(2 / 3)
    assert L.orc_halton(4, 3) == pytest.approx(4 / 9)
    assert L.orc_halton(0, 2) == 0.0


def test_disc_temperature_closed_form(orc):
    L = orc.strict()
    r_in, peak = 6.0, 4500.0
    assert L.orc_disc_temperature_f64(r_in, r_in, peak) == 0.0
    assert L.orc_disc_temperature_f64(5.9, r_in, peak) == 0.0
    rs = np.linspace(r_in, 40 * r_in, 200001)
    T = np.array([L.orc_disc_temperature_f64(float(r), r_in, peak) for r in rs[::50]])
    # max of u^0.75 (1 - sqrt u), u = r_in/r: 0.75 u^-1/4 = 1.25 u^1/4  ->  sqrt(u) = 0.6
    umax = 0.36
    want = peak * umax**0.75 * (1 - math.sqrt(umax)) / 0.186
    assert T.max() == pytest.approx(want, rel=2e-6)
    assert rs[::50][np.argmax(T)] == pytest.approx(r_in / umax, rel=5e-3)
    # far-field slope r^-3/4
    t1 = L.orc_disc_temperature_f64(1e6, 1.0, 1.0)
    t2 = L.orc_disc_temperature_f64(2e6, 1.0, 1.0)
    assert t1 / t2 == pytest.approx(2**0.75, rel=1e-3)


def test_blackbody_properties(orc):
    L = orc.strict()
    c = np.zeros(3)
    L.orc_blackbody_f64(1500.0, p(c))
    assert c[0] > 0.95 and c[1] < c[0] and c[2] < c[1] and c[2] < 0.5
    L.orc_blackbody_f64(6500.0, p(c))
    m = c.mean()
    assert np.all(np.abs(c - m) < 0.15) and m > 0.8
    L.orc_blackbody_f64(30000.0, p(c))
    assert c[2] > 0.95 and c[0] < c[2] and c[1] < c[2] and c[0] < c[1]
    # clamps: below 1000 K and above 40000 K saturate
    a, b = np.zeros(3), np.zeros(3)
    L.orc_blackbody_f64(10.0, p(a))
    L.orc_blackbody_f64(1000.0, p(b))
    np.testing.assert_array_equal(a, b)
    # continuity at the 6600 K branch
    L.orc_blackbody_f64(6599.999, p(a))
    L.orc_blackbody_f64(6600.001, p(b))
    assert np.abs(a - b).max() < 0.02


def test_tone_map_known_answers(orc):
    L = orc.strict()
    o = np.zeros(3, np.float32)
    L.orc_aces_f32(p(np.array([0.0, 1.0, 1e6], np.float32)), p(o))
    assert o[0] == 0.0
    assert o[1] == pytest.approx(2.54 / 3.16, rel=1e-6)  # (2.51 + 0.03) / (2.43 + 0.59 + 0.14)
    assert o[2] == 1.0
    assert L.orc_linear_to_srgb_f32(1.0) == pytest.approx(1.0, abs=1e-6)
    assert L.orc_linear_to_srgb_f32(0.0031308) == pytest.approx(0.04045, abs=2e-5)
    assert L.orc_to_u8_f32(0.5) == 188  # linear 0.5 -> sRGB 0.7354 -> 188
    assert L.orc_to_u8_f32(0.0) == 0 and L.orc_to_u8_f32(1.0) == 255 and L.orc_to_u8_f32(2.0) == 255


def test_starfield_hash_and_density(orc):
    L = orc.strict()
    # xorshift32 (13, 17, 5) of 1 is 270369: mix32(0, 1) = xorshift32(0 * golden + 1)
    assert L.orc_mix32(0, 1) == 270369
    seeds = np.array([L.orc_mix32(ix, iy) for ix in range(0, 768, 3) for iy in range(0, 384, 3)], np.uint64)
    frac = float(np.mean((seeds & 0xFF) < 4))
    assert frac == pytest.approx(4 / 256, abs=0.004)
    c = np.zeros(3, np.float32)
    L.orc_starfield_f32(0.0, 0.0, 0.0, p(c))
    np.testing.assert_array_equal(c, np.array([0.0, 0.0, 0.005], np.float32))


def test_op_counts_are_stable(orc):
    """Algorithmic operation counts per RK4 step used by the roofline."""
    k = orc.count_step(kerr=True)
    s = orc.count_step(kerr=False)
    import singularity_amd.flops as fl

    assert fl.KERR_STEP_OPS == k
    assert fl.SCHW_STEP_OPS == s
