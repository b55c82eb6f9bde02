"""Pins of the oracle's error-controlled Kerr step (step doubling, used for the
16-spp adaptive workload).  The reference has no adaptive integrator, so the
pins are mathematical:

* local error: every accepted step's actual error (against 64 classical RK4
  substeps of the same size, which is exact to ~1e-13 here) stays within 1.5x
  the tolerance in the controller's own norm, for random states and along a
  trajectory; along a trajectory the median sits in [0.1, 1] x tol (the
  controller is neither lax nor wastefully strict);
* the step count scales like tol^(-1/5) for a fourth-order method;
* a whole ray traced at tol 1e-10 converges to the fixed-step answer at
  h = 0.004 M (same class, crossing radius within 5e-6 M: the cubic dense
  output at the crossing; the linear interpolation of the fixed-step
  algorithm would be off by ~3e-3 M on these 6 M steps);
* an escaping ray's exit direction (dense output at r = escape radius) agrees
  with the fine fixed-step one to 3e-6 rad (the fine run's own overshoot of one
  0.004 M step is ~1e-6; taking the first state beyond the radius, as the
  fixed-step algorithm does, would be off by ~1e-3 on 6 M steps);
* the controller halves on rejection (never below 1e-3 h_step), grows 4x after an
  exact step and caps growth at 16 h_step.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import config2, scene_uniforms


def adaptive_step(O, y, c, tol, h_cur, h_step=0.4):
    hc = np.array([h_cur])
    out = np.zeros(5)
    h = O.orc_kerr_adaptive_step(y.ctypes.data, c.ctypes.data, h_step, tol, hc.ctypes.data, out.ctypes.data)
    return out, h, hc[0]


def exact_step(O, y, c, h):
    ref = np.zeros(5)
    O.orc_kerr_rk4_substeps(y.ctypes.data, c.ctypes.data, h, 64, ref.ctypes.data)
    return ref


def norm_err(out, ref):
    d = np.abs(out - ref)
    return max(d[0] / (1 + abs(ref[0])), d[1], d[2])


@pytest.mark.parametrize("tol", [1e-5, 1e-6, 1e-8, 1e-10])
def test_local_error_random_states(orc, tol):
    O = orc.strict()
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(300):
        y = np.array([rng.uniform(3.0, 40.0), rng.uniform(0.2, 2.9), rng.uniform(0, 6), rng.normal(0, 1.5),
                      rng.normal(0, 3)])
        c = np.array([1.0, rng.normal(0, 4), rng.uniform(0, 0.99), 1.0])
        out, h, _ = adaptive_step(O, y, c, tol, 0.4)
        worst = max(worst, norm_err(out, exact_step(O, y, c, h)) / tol)
    assert worst <= 1.5, worst


def trajectory(O, tol, n_max=2000):
    y = np.array([30.0, 1.2, 0.0, -0.9, 1.5])
    c = np.array([1.0, 3.5, 0.9, 1.0])
    hc, ratios = 0.4, []
    for _ in range(n_max):
        out, h, hc = adaptive_step(O, y, c, tol, hc)
        ratios.append(norm_err(out, exact_step(O, y, c, h)) / tol)
        y = out
        if y[0] < 2.0 or y[0] > 60.0:
            break
    return np.array(ratios)


def test_trajectory_error_and_step_scaling(orc):
    O = orc.strict()
    n = {}
    for tol in (1e-6, 1e-8, 1e-10):
        r = trajectory(O, tol)
        n[tol] = len(r)
        steady = r[10:]
        assert steady.max() <= 1.5, (tol, steady.max())
        assert 0.1 <= np.median(steady) <= 1.0, (tol, np.median(steady))
    # fourth order: steps ~ tol^(-1/5); 1e4 in tol -> 10^0.8 = 6.3x the steps
    ratio = n[1e-10] / n[1e-6]
    assert 3.0 <= ratio <= 9.0, (n, ratio)


def test_controller_bounds(orc):
    O = orc.strict()
    y = np.array([40.0, 1.4, 0.0, 0.9, 0.5])
    c = np.array([1.0, 2.0, 0.5, 1.0])
    # generous tolerance in the far field: growth is capped at 16 h_step
    hc = 0.4
    for _ in range(10):
        y, h, hc = adaptive_step(O, y, c, 1e-2, hc)
    assert hc == pytest.approx(16 * 0.4)
    # impossible tolerance: the step halves until the two estimates agree to the
    # last bit (err = 0) or h_cur reaches the floor 1e-3 h_step; then growth is 4x
    y0 = np.array([6.0, 1.0, 0.0, 0.3, 2.0])
    out, h, hc = adaptive_step(O, y0, c, 1e-30, 0.4)
    sin_scale = min(1.0, max(0.05, 5 * abs(np.sin(y0[1]))))
    k = np.log2(0.4 * sin_scale / h)
    assert k == round(k) and 3 <= k <= 10, k
    assert hc == pytest.approx(4 * h / sin_scale)


def test_whole_ray_converges_to_fine_fixed_step(orc):
    u = config2(48, 27, 1)
    fine = scene_uniforms(spin=0.94, r_in=u.disc_r_inner, r_out=25.0, width=48, height=27, dist=40.0,
                          elev=np.radians(10.0), fov=30.0, h=0.004, max_steps=400000)
    ex = orc.Extras()
    ex.adapt_tol = 1e-10
    hits = 0
    for gx, gy in ((16, 0), (12, 3), (20, 3), (16, 9), (12, 15), (20, 21)):
        a = orc.trace_ray(u, gx, gy, precision=64, extras=ex)
        f = orc.trace_ray(fine, gx, gy, precision=64)
        assert a["cls"] == f["cls"], (gx, gy, a, f)
        if a["cls"] == 2:  # disc (oracle class numbering)
            hits += 1
            assert abs(a["r_cross"] - f["r_cross"]) <= 5e-6, (gx, gy, a["r_cross"], f["r_cross"])
        assert a["steps"] < f["steps"] / 5
    assert hits == 6


def test_escape_direction_converges_to_fine_fixed_step(orc):
    u = config2(48, 27, 1)
    fine = scene_uniforms(spin=0.94, r_in=u.disc_r_inner, r_out=25.0, width=48, height=27, dist=40.0,
                          elev=np.radians(10.0), fov=30.0, h=0.004, max_steps=400000)
    ex = orc.Extras()
    ex.adapt_tol = 1e-10
    for gx, gy in ((0, 0), (47, 0), (0, 26), (47, 26)):
        a = orc.trace_ray(u, gx, gy, precision=64, extras=ex)
        f = orc.trace_ray(fine, gx, gy, precision=64)
        assert a["cls"] == f["cls"] == 1  # escape (oracle class numbering)
        assert abs(a["esc_theta"] - f["esc_theta"]) <= 3e-6 and abs(a["esc_phi"] - f["esc_phi"]) <= 3e-6, (a, f)
