/* CPU oracle of the vertical diffusion of buoyancy and velocity with the
 * analytic or TKE-closure eddy coefficients (test infrastructure only: tests,
 * smoke and the bench's CPU baseline use it as the checker). */
#ifndef WTO_VDIFF_ORACLE_H
#define WTO_VDIFF_ORACLE_H
#include "trisk_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* the module variables of vert_diffusion_mod (vert_diffusion.f90:27-64) and
 * the switches of shared.f90 it reads */
typedef struct {
  int tke_closure, patankar, enhance_diff;
  double C_e, C_eps, C_l, C_k, C_srf, e_0, e_min, e_min_srf, eps_s, kappa_VK, Kt_enh, Kt_max, Kt_min, Kv_max,
      Kv_min, mixed_layer, l_0, l_min, Neps_sq, Nsq_min, Q_sr, c_p, bottom_friction;
} wto_vdiff;

/* LAPACK dgtsv with one right-hand side (dgtsv.f90): solves in place,
 * overwriting dl, d, du; returns info (0, or the zero pivot's 1-based row) */
int wto_dgtsv(int n, double* dl, double* d, double* du, double* b);

/* vertical_diffusion (vert_diffusion.f90:66-86) on one uniform level with
 * sol_mean = 0 and no porosity: turbulent_diffusion on the owned nodes and the
 * pole masters (Kt, Kv [L + 1][ncells] at the interfaces, tke [L][ncells];
 * istep 1 recomputes Kt and Kv from the TKE first), then backwards_euler_temp
 * there and backwards_euler_velo on the owned edges, then the ghost fill.
 * Forcing per node: bottom and top buoyancy fluxes, wind-stress magnitude
 * [ncells], wind flux [3][ncells]; eta NULL: the free surface of the column.
 * The state must be ghost filled on entry. */
void wto_vertical_diffusion(const wto_grid* g, const wto_vdiff* p, int L, double dt, int istep, const double* bot,
                            const double* top, const double* tau, const double* wind, const double* eta, double* Kt,
                            double* Kv, double* tke, double* mass, double* temp, double* velo);

#ifdef __cplusplus
}
#endif
#endif
