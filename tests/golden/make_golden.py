"""Generate tests/golden/ref_vectors.npz from the reference's own headers.

Runs only where /root/reference exists: it builds oracle/_ref/libref.so (the
reference's shared_shader/*.h and core/include/physics/kerr.hpp compiled where
they lie, see oracle/Makefile) and records input/output vectors of:

  * kerr_ham_rhs on the verification/test_kerr_ham_analytic.py grid
    (r x theta x a x L_z with E=1, M=1, p_r=0.6, p_theta=2.3),
  * kerr_ham_rk4_step trajectories (2000 steps, every 25th state kept) at
    a/M in {0, 0.5, 0.94, 0.998},
  * geodesic_rhs_schwarzschild / rk4_step trajectories,
  * detect_equatorial_crossing on random pairs,
  * Kerr closed forms (horizons, ISCO, photon orbits) on a spin grid,
  * the Uniforms byte offsets.

With the argument `shading` it writes tests/golden/ref_shading.npz instead:
the shading arithmetic's host twins in core/include/physics/disc.hpp:122-155
(blackbody_srgb_tanner_helland) and redshift.hpp:44-112
(schwarzschild_grav_redshift, _keplerian_omega, _doppler_factor,
intensity_scaling) on temperature, radius, cos(psi) and g grids.

Usage: python tests/golden/make_golden.py [shading]
"""

from __future__ import annotations

import ctypes as C
import itertools
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402


def f32p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def kerr_init_state(a: float, r0: float, theta0: float, ur: float, uth: float, uph: float):
    """Null initial data at (r0, theta0) from a coordinate velocity, in float64
    (the vectors only need valid photon states; their values are recorded)."""
    M = 1.0
    ct, st = np.cos(theta0), np.sin(theta0)
    Sig = r0 * r0 + a * a * ct * ct
    Del = r0 * r0 - 2 * M * r0 + a * a
    s2 = st * st
    g_tt = -(1 - 2 * M * r0 / Sig)
    g_tp = -2 * M * r0 * a * s2 / Sig
    g_rr = Sig / Del
    g_pp = s2 * ((r0 * r0 + a * a) ** 2 - a * a * Del * s2) / Sig
    A, B, Cq = g_tt, 2 * g_tp * uph, g_rr * ur * ur + Sig * uth * uth + g_pp * uph * uph
    ut = (-B - np.sqrt(B * B - 4 * A * Cq)) / (2 * A)
    E = -(g_tt * ut + g_tp * uph)
    L = g_tp * ut + g_pp * uph
    pth = Sig * uth
    Q = pth * pth + ct * ct * (L * L / max(s2, 1e-6) - a * a * E * E)
    state = np.array([0.0, r0, theta0, 0.0, g_rr * ur, pth], np.float32)
    cons = np.array([E, L, Q, a, M], np.float32)
    return state, cons


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    if not oracle.ref_available():
        raise SystemExit("reference tree not available: cannot generate vectors")
    R = oracle.ref()
    R.ref_kerr_rhs.argtypes = [C.c_void_p] * 3
    R.ref_kerr_two_H.restype = C.c_float
    R.ref_kerr_two_H.argtypes = [C.c_void_p] * 2
    R.ref_kerr_rk4.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_void_p]
    R.ref_schw_rhs.argtypes = [C.c_void_p, C.c_float, C.c_void_p]
    R.ref_schw_rk4.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_void_p]
    R.ref_crossing.argtypes = [C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    R.ref_kerr_closed_form.restype = C.c_float
    R.ref_kerr_closed_form.argtypes = [C.c_int, C.c_float, C.c_float]
    R.ref_uniform_offsets.restype = C.c_int
    R.ref_uniform_offsets.argtypes = [C.c_void_p]
    R.ref_sizeof_uniforms.restype = C.c_int
    out: dict[str, np.ndarray] = {}

    # kerr_ham_rhs on the analytic-partials grid
    grid = list(itertools.product([2.6, 3.1, 6.0, 12.0, 40.0, 150.0],
                                  [0.05, 0.4, 1.0, np.pi / 2, 2.4, np.pi - 0.05],
                                  [0.0, 0.5, 0.94, 0.998], [-7.5, 0.0, 3.7]))
    S = np.zeros((len(grid), 6), np.float32)
    K = np.zeros((len(grid), 5), np.float32)
    D = np.zeros((len(grid), 6), np.float32)
    H2 = np.zeros(len(grid), np.float32)
    for i, (r, th, a, L) in enumerate(grid):
        S[i] = [0.0, r, th, 0.3, 0.6, 2.3]
        K[i] = [1.0, L, 0.0, a, 1.0]
        R.ref_kerr_rhs(f32p(S[i]), f32p(K[i]), f32p(D[i]))
        H2[i] = R.ref_kerr_two_H(f32p(S[i]), f32p(K[i]))
    out.update(kerr_rhs_state=S, kerr_rhs_const=K, kerr_rhs_out=D, kerr_two_H=H2)

    # kerr rk4 trajectories
    trajs, starts, consts, hs = [], [], [], []
    rng = np.random.default_rng(1234)
    for a in (0.0, 0.5, 0.94, 0.998):
        for _ in range(3):
            th0 = float(rng.uniform(0.6, 2.5))
            ur = -float(rng.uniform(0.5, 1.0))
            uth = float(rng.uniform(-0.02, 0.02))
            uph = float(rng.uniform(-0.02, 0.02))
            st, cs = kerr_init_state(a, 30.0, th0, ur, uth, uph)
            n, h = 2000, np.float32(0.2)
            tr = np.zeros((n + 1, 6), np.float32)
            R.ref_kerr_rk4(f32p(st), h, f32p(cs), n, f32p(tr))
            trajs.append(tr)
            starts.append(st)
            consts.append(cs)
            hs.append(h)
    out.update(kerr_rk4_start=np.array(starts), kerr_rk4_const=np.array(consts),
               kerr_rk4_h=np.array(hs, np.float32), kerr_rk4_traj=np.array(trajs)[:, ::25])

    # schwarzschild rhs + rk4
    ss, sd, st_all, tr_all = [], [], [], []
    for _ in range(40):
        s = np.array([0, rng.uniform(2.5, 40), rng.uniform(0.2, 2.9), rng.uniform(0, 6), rng.uniform(1, 2),
                      rng.uniform(-1, 1), rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)], np.float32)
        d = np.zeros(8, np.float32)
        R.ref_schw_rhs(f32p(s), np.float32(2.0), f32p(d))
        ss.append(s)
        sd.append(d)
    for _ in range(6):
        r0 = 30.0
        s = np.array([0, r0, rng.uniform(1.0, 2.1), 0, 0, -1.0, rng.uniform(-0.004, 0.004),
                      rng.uniform(-0.004, 0.004)], np.float32)
        f = 1 - 2.0 / r0
        kin = s[5] ** 2 / f + (s[6] * r0) ** 2 + (s[7] * r0 * np.sin(s[2])) ** 2
        s[4] = np.float32(np.sqrt(kin / f))
        tr = np.zeros((1501, 8), np.float32)
        R.ref_schw_rk4(f32p(s), np.float32(0.1), np.float32(2.0), 1500, f32p(tr))
        st_all.append(s)
        tr_all.append(tr)
    out.update(schw_rhs_state=np.array(ss), schw_rhs_out=np.array(sd), schw_rk4_start=np.array(st_all),
               schw_rk4_traj=np.array(tr_all)[:, ::25])

    # crossings
    pairs = rng.uniform(1.2, 1.95, size=(400, 2)).astype(np.float32)
    crossed = np.zeros(400, np.int32)
    frac = np.zeros(400, np.float32)
    for i, (p, q) in enumerate(pairs):
        c, fr = C.c_int(0), C.c_float(0)
        R.ref_crossing(float(p), float(q), C.byref(c), C.byref(fr))
        crossed[i], frac[i] = c.value, fr.value
    out.update(cross_pairs=pairs, cross_crossed=crossed, cross_frac=frac)

    # closed forms
    spins = np.array([0.0, 0.1, 0.3, 0.5, 0.7, 0.9, 0.94, 0.99, 0.998], np.float32)
    cf = np.array([[R.ref_kerr_closed_form(w, 1.0, float(a)) for w in range(6)] for a in spins], np.float32)
    out.update(closed_spins=spins, closed_forms=cf)

    offs = np.zeros(64, np.int32)
    n = R.ref_uniform_offsets(f32p(offs))
    out.update(uniform_offsets=offs[:n], uniform_size=np.array([R.ref_sizeof_uniforms()], np.int32))
    np.savez_compressed(HERE / "ref_vectors.npz", **out)
    print("wrote", HERE / "ref_vectors.npz", {k: v.shape for k, v in out.items()})


def shading() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    if not oracle.ref_available():
        raise SystemExit("reference tree not available: cannot generate vectors")
    R = oracle.ref()
    R.ref_blackbody_th.argtypes = [C.c_float, C.c_void_p]
    for name in ("ref_grav_redshift", "ref_doppler_factor", "ref_keplerian_omega", "ref_intensity_scaling"):
        getattr(R, name).restype = C.c_float
    R.ref_grav_redshift.argtypes = [C.c_float] * 3
    R.ref_doppler_factor.argtypes = [C.c_float] * 4
    R.ref_keplerian_omega.argtypes = [C.c_float] * 2
    R.ref_intensity_scaling.argtypes = [C.c_float]
    out: dict[str, np.ndarray] = {}
    T = np.unique(np.concatenate([np.geomspace(1000.0, 40000.0, 4001), np.linspace(1000.0, 40000.0, 391),
                                  [6599.0, 6600.0, 6601.0, 1900.0, 1901.0]]).astype(np.float32))
    rgb = np.zeros((T.size, 3), np.float32)
    for i, t in enumerate(T):
        R.ref_blackbody_th(float(t), f32p(rgb[i]))
    out.update(bb_T=T, bb_rgb=rgb)
    # Schwarzschild (M = 1, rs = 2): emitter radii outside the horizon, observers at 30 and 100 M
    r = np.geomspace(2.05, 60.0, 64).astype(np.float32)
    r_obs = np.array([30.0, 100.0], np.float32)
    cpsi = np.linspace(-1.0, 1.0, 21).astype(np.float32)
    grav = np.array([[R.ref_grav_redshift(2.0, float(x), float(o)) for o in r_obs] for x in r], np.float32)
    omega = np.array([R.ref_keplerian_omega(1.0, float(x)) for x in r], np.float32)
    dop = np.array([[R.ref_doppler_factor(2.0, float(x), float(w), float(c)) for c in cpsi]
                    for x, w in zip(r, omega)], np.float32)
    g = np.geomspace(0.05, 3.0, 97).astype(np.float32)
    ints = np.array([R.ref_intensity_scaling(float(v)) for v in g], np.float32)
    out.update(rs_r=r, rs_r_obs=r_obs, rs_cos_psi=cpsi, rs_grav=grav, rs_omega=omega, rs_doppler=dop,
               rs_g=g, rs_intensity=ints)
    np.savez_compressed(HERE / "ref_shading.npz", **out)
    print("wrote", HERE / "ref_shading.npz", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    if sys.argv[1:] == ["shading"]:
        shading()
    else:
        main()
