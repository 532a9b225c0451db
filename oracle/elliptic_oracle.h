/* CPU oracle of the free-surface elliptic path (see elliptic_oracle.c).
 * TEST INFRASTRUCTURE ONLY. */
#ifndef WTG_ELLIPTIC_ORACLE_H
#define WTG_ELLIPTIC_ORACLE_H
#include "trisk_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double grav, dt, theta1, theta2;
} wto_fs_params;

/* owner copy of a node field into its halo cells */
void wto_node_ghost(const wto_grid* g, double* f);
/* Lq = theta1 theta2 dt^2 div((H + eta^n) g grad q) - q; inputs ghost filled, output ghost filled */
void wto_elliptic_apply(const wto_grid* g, const wto_fs_params* p, const double* eta_old, const double* q, double* Lq);
void wto_elliptic_diag(const wto_grid* g, const wto_fs_params* p, const double* eta_old, double* diag);
/* u += w (f - L u) / diag */
void wto_jacobi(const wto_grid* g, const wto_fs_params* p, const double* eta_old, double* u, const double* f, double w);
/* sum_c wgt[c] a[c] b[c] */
double wto_wdot(const wto_grid* g, const double* wgt, const double* a, const double* b);
/* returns the iteration count (maxit + 1 if tol was not reached) */
int wto_bicgstab(const wto_grid* g, const wto_fs_params* p, const double* eta_old, const double* wgt, double* u,
                 const double* f, double nrm_f, double tol, int maxit, double* err_out);
/* divergence of the layer-integrated mass flux sum_k u_k interp(mass_k) pedlen */
void wto_flux_divergence(const wto_grid* g, int L, const double* mass, const double* velo, double* divF);
/* the same, also returning the layer-integrated edge fluxes [3][ncells] of
 * the four cells around every master (0 elsewhere) */
void wto_flux_divergence_parts(const wto_grid* g, int L, const double* mass, const double* velo, double* divF,
                               double* flux_out);
/* rhs = -eta + dt (theta2 divF + (1 - theta2) divF_old) / rho0 */
void wto_rhs_elliptic(const wto_grid* g, const wto_fs_params* p, double rho0, const double* eta, const double* divF,
                      const double* divF_old, double* rhs);

#ifdef __cplusplus
}
#endif
#endif
