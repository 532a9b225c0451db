/* CPU oracle of the split-explicit (mode split) time step (see split_oracle.c).
 * TEST INFRASTRUCTURE ONLY. */
#ifndef WTG_SPLIT_ORACLE_H
#define WTG_SPLIT_ORACLE_H
#include "elliptic_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* layer dilation of mass / temp to the free surface eta (masters, then ghosts) */
void wto_barotropic_correction(const wto_grid* g, int L, const double* eta, double* mass, double* temp);
/* g grad(eta) on the three edges of every cell whose edges lie in the tile
 * ([3][ncells], 0 elsewhere) */
void wto_grad_eta(const wto_grid* g, const double* eta, double* hflux);
/* one RK4_split step of size dt on (mass, temp, velo, eta), the free surface
 * solved by BiCGSTAB to tol (maxit iterations); returns its iteration count */
int wto_rk4_split(const wto_grid* g, const wto_fs_params* p, const double* wgt, int L, double tol, int maxit,
                  double* mass, double* temp, double* velo, double* eta);
/* the same with nstages = 1..4 stages: Euler_split, RK2_split, RK3_split,
 * RK4_split (stage coefficients dt/nstages, ..., dt/2, dt) */
int wto_rk_split(const wto_grid* g, const wto_fs_params* p, const double* wgt, int L, int nstages, double tol,
                 int maxit, double* mass, double* temp, double* velo, double* eta);
/* the same with the trend's Laplacian diffusion / hyperdiffusion (trend_ml
 * calls it in every stage of the split schemes too) */
int wto_rk_split_diff(const wto_grid* g, const wto_fs_params* p, const wto_diff* dif, const double* wgt, int L,
                      int nstages, double tol, int maxit, double* mass, double* temp, double* velo, double* eta);

#ifdef __cplusplus
}
#endif
#endif
