/* include/sing_diag.h — diagnostic entry points of libsing_hip.so.
 *
 * No counterpart in the reference.  They evaluate, on the host, the same
 * templated code the adjoint kernels run on the device (Kerr RHS, damped RK4
 * step and their hand-written vector-Jacobian products), so unit tests can pin
 * the discrete adjoint against exact Jacobians without a GPU.  Double
 * precision; y = (r, theta, phi, p_r, p_theta), c = (E, L_z, a, M).
 */
#ifndef SING_DIAG_H
#define SING_DIAG_H

#ifdef __cplusplus
extern "C" {
#endif

void sing_diag_kerr_rhs(const double* y5, const double* c4, double* out5);
/* gy5 = v^T d rhs/dy, gc3 = v^T d rhs/d(E, L_z, a) */
void sing_diag_kerr_rhs_vjp(const double* y5, const double* c4, const double* v5, double* gy5, double* gc3);
void sing_diag_kerr_step(const double* y5, const double* c4, double h, double* out5);
/* gy5 = lam^T d step/dy, gc3 = lam^T d step/d(E, L_z, a) */
void sing_diag_kerr_step_vjp(const double* y5, const double* c4, double h, const double* lam5, double* gy5,
                             double* gc3);

#ifdef __cplusplus
}
#endif

#endif /* SING_DIAG_H */
