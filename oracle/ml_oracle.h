/* CPU oracle of the multi-level explicit step on a full hierarchy of uniform
 * levels (see ml_oracle.c).  TEST INFRASTRUCTURE ONLY. */
#ifndef WTG_ML_ORACLE_H
#define WTG_ML_ORACLE_H
#include "mg_oracle.h"
#include "trisk_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* levels 0 .. nlevels-1, coarse to fine.  Inter-level data of the pair
 * (l-1, l) lives with the finer level l (ovl, split, iu, wide) or the coarser
 * one (rf): ovl [4][NC_l], split [2][NC_l], iu [9][NC_l], rf [24][NC_{l-1}],
 * wide_src/WH: the fine level's owned cell of lattice points up to WH cells
 * beyond its owned square (frame [-WH, n+WH)^2). */
typedef struct {
  int nlevels;
  const wto_grid* grid[16];
  const double* ovl[16];
  const double* split[16];
  const double* iu[16];
  const double* rf[16];
  const int64_t* wide_src[16];
  int WH[16];
} wto_ml;

/* velocity wavelets of level l from levels l and l-1 (Compute_velo_wavelets
 * and its pentagon corrections), one layer: wc [3][NC_l], ghost filled */
void wto_velo_wavelets(const wto_ml* h, int l, const double* u_f, const double* u_c, double* wc);
/* velocity of level l from level l-1 and the wavelets (Reconstruct_outer_velo,
 * Reconstruct_velo_penta, Reconstruct_inner_velo), one layer, ghost filled */
void wto_velo_reconstruct(const wto_ml* h, int l, const double* u_c, const double* wc, double* u_f);
/* flux of level l-1 from the flux and scalar trend of level l (flux_cpt_restr),
 * one scalar field of one layer: hc [3][NC_{l-1}], ghost filled */
void wto_flux_restrict(const wto_ml* h, int l, const double* hf, const double* df, double* hc);

/* trend_ml on the hierarchy (every mask TOLRNZ): state and trend per level,
 * [L][cell] scalars, [L][3][cell] velocity; inputs ghost filled */
void wto_trend_ml(const wto_ml* h, int L, double* const* mass, double* const* temp, double* const* velo,
                  double** dmass, double** dtemp, double** dvelo);
/* WT_after_step (adapt.f90:192-237) in place on every level */
void wto_wt_after_step(const wto_ml* h, int L, double** mass, double** temp, double** velo);
/* nsteps RK4 steps (time_integr.f90:78-103) on the hierarchy, in place */
void wto_rk4_ml(const wto_ml* h, int L, double dt, int nsteps, double** mass, double** temp, double** velo);

/* one RK4_split step (time_integr.f90:360-382, 444-478) on the hierarchy:
 * per stage barotropic_correction on every level, trend_ml, scalar_star and
 * u_star with the external pressure gradient of every level (coarse eta
 * sub-sampled), WT_after_step; then flux_divergence over levels, rhs_elliptic,
 * FMG (mg: the same levels, its dt/theta/solver parameters; its eta_old
 * pointers are set here), barotropic_correction, u_update and WT_after_step
 * with the free surface.  eta[l] per level; fmg_iters[l] = FMG's counts. */
int wto_rk4_split_ml(const wto_ml* h, wto_mg* mg, int L, double** mass, double** temp, double** velo,
                     double** eta, int* fmg_iters);
/* the same with nstages = 1..4 (Euler_split, RK2_split, RK3_split, RK4_split) */
int wto_rk_split_ml(const wto_ml* h, wto_mg* mg, int L, int nstages, double** mass, double** temp, double** velo,
                    double** eta, int* fmg_iters);

#ifdef __cplusplus
}
#endif
#endif
