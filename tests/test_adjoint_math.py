"""The hand-written discrete adjoint, evaluated on the host through the
diagnostics ABI (include/sing_diag.h: the same templated code the adjoint
kernels run on the GPU), against exact Jacobians.

* kerr_rhs / kerr_step equal the oracle's restatement of the reference
  (shared_shader/kerr_hamilton.h:97-178) to rounding.
* kerr_step_vjp(lambda) equals J^T lambda with J the oracle's forward-mode
  dual-number Jacobian of the reference RK4 step (independent code:
  reverse-mode by hand vs dual numbers), including the pole-damping branch.
* kerr_rhs_vjp agrees with central differences of kerr_rhs.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import singularity_amd as sa

P = C.c_void_p


def lib():
    L = sa.lib()
    L.sing_diag_kerr_rhs.argtypes = [P, P, P]
    L.sing_diag_kerr_rhs_vjp.argtypes = [P, P, P, P, P]
    L.sing_diag_kerr_step.argtypes = [P, P, C.c_double, P]
    L.sing_diag_kerr_step_vjp.argtypes = [P, P, C.c_double, P, P, P]
    return L


def ptr(a):
    return a.ctypes.data_as(P)


def random_states(n, seed, pole=False):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        th = rng.uniform(0.002, 0.2) if pole else rng.uniform(0.3, 2.8)
        if pole and rng.random() < 0.5:
            th = np.pi - th
        y = np.array([rng.uniform(2.5, 40.0), th, rng.uniform(0, 6), rng.normal(0, 2), rng.normal(0, 3)])
        c = np.array([rng.uniform(0.7, 1.3), rng.normal(0, 4), rng.uniform(0.0, 0.99), 1.0])  # E, L, a, M
        out.append((y, c))
    return out


def test_rhs_and_step_match_oracle(orc):
    L, O = lib(), orc.strict()
    got, ref = np.zeros(5), np.zeros(6)
    for y, c in random_states(300, 1):
        L.sing_diag_kerr_rhs(ptr(y), ptr(c), ptr(got))
        s6 = np.array([0.0, *y])
        c5 = np.array([c[0], c[1], 0.0, c[2], c[3]])
        O.orc_kerr_rhs_f64(ptr(s6), ptr(c5), ptr(ref))
        np.testing.assert_allclose(got, ref[1:], rtol=1e-11, atol=1e-13)
        tr = np.zeros((2, 6))
        O.orc_kerr_rk4_f64(ptr(s6), 0.4, ptr(c5), 1, ptr(tr))
        L.sing_diag_kerr_step(ptr(y), ptr(c), 0.4, ptr(got))
        np.testing.assert_allclose(got, tr[1, 1:], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("pole", [False, True])
def test_step_vjp_equals_exact_jacobian_transpose(orc, pole):
    L, O = lib(), orc.strict()
    O.orc_kerr_step_jacobian.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, P]
    rng = np.random.default_rng(7 if pole else 8)
    gy, gc = np.zeros(5), np.zeros(3)
    J = np.zeros((5, 8))
    for y, c in random_states(200, 3 if pole else 4, pole):
        for h in (0.4, 0.05):
            lam = rng.normal(size=5)
            L.sing_diag_kerr_step_vjp(ptr(y), ptr(c), h, ptr(lam), ptr(gy), ptr(gc))
            O.orc_kerr_step_jacobian(ptr(y), c[0], c[1], c[2], c[3], h, ptr(J))
            want = J.T @ lam  # d/d(r, theta, phi, p_r, p_theta, E, L, a)
            got = np.concatenate([gy, gc])
            scale = np.abs(J.T).dot(np.abs(lam)) + 1e-12
            assert np.all(np.abs(got - want) <= 1e-9 * scale + 1e-12), (y, c, h, got, want)


def test_rhs_vjp_matches_central_differences():
    L = lib()
    rng = np.random.default_rng(11)
    gy, gc = np.zeros(5), np.zeros(3)
    f_p, f_m = np.zeros(5), np.zeros(5)
    for y, c in random_states(100, 12):
        v = rng.normal(size=5)
        L.sing_diag_kerr_rhs_vjp(ptr(y), ptr(c), ptr(v), ptr(gy), ptr(gc))
        x = np.concatenate([y, c[:3]])
        fd = np.zeros(8)
        for j in range(8):
            e = 1e-6 * max(1.0, abs(x[j]))
            xp, xm = x.copy(), x.copy()
            xp[j] += e
            xm[j] -= e
            L.sing_diag_kerr_rhs(ptr(np.ascontiguousarray(xp[:5])), ptr(np.array([*xp[5:], c[3]])), ptr(f_p))
            L.sing_diag_kerr_rhs(ptr(np.ascontiguousarray(xm[:5])), ptr(np.array([*xm[5:], c[3]])), ptr(f_m))
            fd[j] = v.dot(f_p - f_m) / (2 * e)
        got = np.concatenate([gy, gc])
        np.testing.assert_allclose(got, fd, rtol=2e-5, atol=1e-7 * (1 + np.abs(fd).max()))
