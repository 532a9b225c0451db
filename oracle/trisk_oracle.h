/* CPU oracle of the TRiSK hot path on the lozenge-tile layout.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * algorithm, used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker.  Nothing in the product links or calls it.
 *
 * Layout (shared with the product): 10 lozenge tiles of S x S cells, S = n + 2*2,
 * cell (lz, i, j) = lz*S*S + (j+2)*S + (i+2), i, j in [-2, n+2); scalar fields
 * are [layer][cell], edge fields [layer][edge RT|DG|UP][cell], geometry planes
 * [component][cell].
 */
#ifndef WTG_TRISK_ORACLE_H
#define WTG_TRISK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n, S;                         /* lozenge side, tile side */
  int64_t ncells;                   /* 10*S*S */
  int64_t south_pole, north_pole;   /* master cells of the poles */
  const double *len, *pedlen;       /* [3][ncells] */
  const double *area;               /* [6][ncells] hexagon part areas */
  const double *hex_inv;            /* [ncells] */
  const double *triarea, *coriolis; /* [2][ncells] */
  const double *topo;               /* [ncells] bottom height z_s (m) */
  const int64_t *node_src;          /* [ncells] */
  const int64_t *edge_src;          /* [3*ncells], slot = 3*cell + e */
  const int8_t *edge_sign;          /* [3*ncells] */
  double grav, rho0, p_top;
} wto_grid;

/* halo copy of L layers of node fields and edge fields (comm.f90:785-861) */
void wto_ghost_fill(const wto_grid* g, int L, double* mass, double* temp, double* velo);

/* trend_ml (multi_level.f90:5-47) on one uniform level, incompressible
 * (Boussinesq) layers, no diffusion, zero mean state.  Inputs must be ghost
 * filled.  Outputs are written on owned cells and the pole masters; other
 * cells are zeroed. */
void wto_trend(const wto_grid* g, int L, const double* mass, const double* temp, const double* velo,
               double* dmass, double* dtemp, double* dvelo);

/* wto_trend with the quantities the coarser levels of a hierarchy restrict
 * (multi_level.f90:49-188), per layer: the mass and temperature fluxes
 * h_flux [L][3][ncells] on every edge the divergence reads, Bernoulli and
 * Exner [L][ncells] on [0, n]^2, and the velocity source (-Qperp, before
 * du_grad divides by the length) [L][3][ncells] on owned edges. */
typedef struct {
  double *hflux_m, *hflux_t, *bern, *exner, *du_src;
} wto_trend_keep;
void wto_trend_parts(const wto_grid* g, int L, const double* mass, const double* temp, const double* velo,
                     double* dmass, double* dtemp, double* dvelo, wto_trend_keep* keep);

/* Laplacian diffusion (order 1) or iterated hyperdiffusion (order 2) of the
 * scalars (lsclr), div u (ldivu) and rot u (lrotu) through the physics
 * hooks in the form of the reference's test cases; the velocity term needs
 * lrotu != 0 (its sign is (-1)^(lrotu-1)) */
typedef struct {
  int lsclr, ldivu, lrotu;
  double visc_m, visc_t, visc_divu, visc_rotu;
  /* test-only planted mistake, 0 in every real use: 1 gives the halo of divu
   * its owners' values and computes the Laplacians' halo fluxes locally */
  int plant_owner_halo;
} wto_diff;
void wto_trend_diff(const wto_grid* g, const wto_diff* dif, int L, const double* mass, const double* temp,
                    const double* velo, double* dmass, double* dtemp, double* dvelo);
void wto_rk4_diff(const wto_grid* g, const wto_diff* dif, int L, double dt, int nsteps, double* mass, double* temp,
                  double* velo);

/* nsteps low-storage RK4 steps (time_integr.f90:78-103, 206-225) in place;
 * the state is ghost filled on entry and on exit. */
void wto_rk4(const wto_grid* g, int L, double dt, int nsteps, double* mass, double* temp, double* velo);

/* threads the oracle's cell loops use (OpenMP; 1 without it) */
int wto_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
