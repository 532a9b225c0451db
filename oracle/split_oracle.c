/* CPU oracle of the split-explicit time step with implicit free surface.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * algorithms, used by tests/ as the checker.  Nothing in the product links
 * or calls it.
 *
 * Restated routines (reference src/):
 *   barotropic_correction, total_height, cal_height     barotropic_2d.f90:52-116, 368-406
 *   grad_eta (step1 itype 3)                            barotropic_2d.f90:457-474, ops.f90:217-220, 418-421
 *   scalar_star, u_star, u_update                        barotropic_2d.f90:9-50, 131-149
 *   RK4_split, RK_split, free_surface_update, eta_update time_integr.f90:360-382, 444-478,
 *                                                        barotropic_2d.f90:118-129
 * Porosity is 1 and the mean state 0 (sol_mean = 0), so phi_node = 1 and
 * porous_density = ref_density.  The elliptic solver is SRJ on one level,
 * i.e. BiCGSTAB to coarse_tol (lin_solve.f90:142-196).  The velocity stage
 * update keeps the reference's form sol + h (trend + theta1 g grad eta).
 * Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "split_oracle.h"

#define H 2

static int in_tile(const wto_grid* g, int64_t c) {
  int64_t r = c % ((int64_t)g->S * g->S);
  int j = (int)(r / g->S) - H, i = (int)(r % g->S) - H;
  return i >= 0 && i < g->n && j >= 0 && j < g->n;
}

static int is_master(const wto_grid* g, int64_t c) {
  return in_tile(g, c) || c == g->south_pole || c == g->north_pole;
}

/* cells with three edges inside the tile (the last row / column has none beyond it) */
static int edges_in_tile(const wto_grid* g, int64_t c) {
  int64_t r = c % ((int64_t)g->S * g->S);
  int j = (int)(r / g->S) - H, i = (int)(r % g->S) - H;
  return i >= -H && i < g->n + H - 1 && j >= -H && j < g->n + H - 1;
}

void wto_barotropic_correction(const wto_grid* g, int L, const double* eta, double* mass, double* temp) {
  const int64_t NC = g->ncells;
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master(g, c)) continue;
    double h = 0.0; /* total_height: sum of layer thicknesses */
    for (int k = 0; k < L; ++k) {
      double rho_dz = 0.0 + mass[(int64_t)k * NC + c];
      h = h + rho_dz / (g->rho0 * 1.0);
    }
    for (int k = 0; k < L; ++k) {
      const int64_t x = (int64_t)k * NC + c;
      double rho_dz = 0.0 + mass[x];
      double rho_dz_theta = 0.0 + temp[x];
      double theta = rho_dz_theta / rho_dz;
      double e = eta[c] / 1.0;
      mass[x] = (e - g->topo[c]) / h * rho_dz - 0.0;
      rho_dz = 0.0 + mass[x];
      temp[x] = rho_dz * theta - 0.0;
    }
  }
  for (int k = 0; k < L; ++k) {
    wto_node_ghost(g, mass + (int64_t)k * NC);
    wto_node_ghost(g, temp + (int64_t)k * NC);
  }
}

void wto_grad_eta(const wto_grid* g, const double* eta, double* hf) {
  const int64_t NC = g->ncells, S = g->S;
  for (int64_t c = 0; c < NC; ++c) {
    if (!edges_in_tile(g, c)) {
      hf[c] = hf[NC + c] = hf[2 * NC + c] = 0.0;
      continue;
    }
    hf[c] = g->grav * (eta[c + 1] / 1.0 - eta[c] / 1.0) / g->len[c];
    hf[NC + c] = g->grav * (eta[c] / 1.0 - eta[c + S + 1] / 1.0) / g->len[NC + c];
    hf[2 * NC + c] = g->grav * (eta[c + S] / 1.0 - eta[c] / 1.0) / g->len[2 * NC + c];
  }
}

/* Euler_split, RK2_split, RK3_split, RK4_split (time_integr.f90:360-451): the
 * low-storage stages RK_split(h/(n-k)) from sol through q1 back to sol, then
 * free_surface_update */
static int rk_split_impl(const wto_grid* g, const wto_fs_params* p, const wto_diff* dif, const double* wgt, int L,
                         int nstages, double tol, int maxit, double* mass, double* temp, double* velo, double* eta);

int wto_rk_split(const wto_grid* g, const wto_fs_params* p, const double* wgt, int L, int nstages, double tol,
                 int maxit, double* mass, double* temp, double* velo, double* eta) {
  return rk_split_impl(g, p, NULL, wgt, L, nstages, tol, maxit, mass, temp, velo, eta);
}

int wto_rk_split_diff(const wto_grid* g, const wto_fs_params* p, const wto_diff* dif, const double* wgt, int L,
                      int nstages, double tol, int maxit, double* mass, double* temp, double* velo, double* eta) {
  return rk_split_impl(g, p, dif, wgt, L, nstages, tol, maxit, mass, temp, velo, eta);
}

static int rk_split_impl(const wto_grid* g, const wto_fs_params* p, const wto_diff* dif, const double* wgt, int L,
                         int nstages, double tol, int maxit, double* mass, double* temp, double* velo, double* eta) {
  const int64_t NC = g->ncells, ns = NC * L, ne = 3 * NC * L;
  double *m1 = malloc(8 * ns), *t1 = malloc(8 * ns), *u1 = malloc(8 * ne);
  double *dm = malloc(8 * ns), *dtm = malloc(8 * ns), *du = malloc(8 * ne);
  double *hf = malloc(8 * 3 * NC), *dF = calloc(NC, 8), *dFold = calloc(NC, 8), *rhs = calloc(NC, 8),
         *eo = malloc(8 * NC);
  wto_ghost_fill(g, L, mass, temp, velo);
  wto_node_ghost(g, eta);
  if (p->theta2 != 1.0) wto_flux_divergence(g, L, mass, velo, dFold);
  memcpy(m1, mass, 8 * ns);
  memcpy(t1, temp, 8 * ns);
  memcpy(u1, velo, 8 * ne);
  wto_grad_eta(g, eta, hf); /* eta is fixed through the stages */
  for (int st = 0; st < nstages; ++st) {
    double* ym = st == 0 ? mass : m1;
    double* yt = st == 0 ? temp : t1;
    double* yu = st == 0 ? velo : u1;
    wto_barotropic_correction(g, L, eta, ym, yt);
    if (dif)
      wto_trend_diff(g, dif, L, ym, yt, yu, dm, dtm, du);
    else
      wto_trend(g, L, ym, yt, yu, dm, dtm, du);
    const int last = st == nstages - 1;
    double* om = last ? mass : m1;
    double* ot = last ? temp : t1;
    double* ou = last ? velo : u1;
    const double h = p->dt / (double)(nstages - st); /* h/4, h/3, h/2, h */
    for (int k = 0; k < L; ++k)
      for (int64_t c = 0; c < NC; ++c) {
        if (!is_master(g, c)) continue;
        const int64_t x = (int64_t)k * NC + c;
        om[x] = mass[x] + h * dm[x];
        ot[x] = temp[x] + h * dtm[x];
        for (int e = 0; e < 3; ++e) {
          const int64_t y = ((int64_t)k * 3 + e) * NC + c;
          ou[y] = velo[y] + h * (du[y] + p->theta1 * hf[e * NC + c]);
        }
      }
    wto_ghost_fill(g, L, om, ot, ou);
  }
  /* free_surface_update */
  wto_flux_divergence(g, L, mass, velo, dF);
  wto_rhs_elliptic(g, p, g->rho0, eta, dF, dFold, rhs);
  memcpy(eo, eta, 8 * NC);
  double nrm = sqrt(wto_wdot(g, wgt, rhs, rhs));
  if (nrm == 0.0) nrm = 1.0;
  double err;
  const int iters = wto_bicgstab(g, p, eo, wgt, eta, rhs, nrm, tol, maxit, &err);
  wto_barotropic_correction(g, L, eta, mass, temp);
  wto_grad_eta(g, eta, hf);
  for (int k = 0; k < L; ++k)
    for (int64_t c = 0; c < NC; ++c) {
      if (!is_master(g, c)) continue;
      for (int e = 0; e < 3; ++e) {
        const int64_t y = ((int64_t)k * 3 + e) * NC + c;
        velo[y] = velo[y] - p->theta1 * p->dt * hf[e * NC + c];
      }
    }
  wto_ghost_fill(g, L, mass, temp, velo);
  free(m1); free(t1); free(u1); free(dm); free(dtm); free(du); free(hf); free(dF); free(dFold); free(rhs); free(eo);
  return iters;
}

int wto_rk4_split(const wto_grid* g, const wto_fs_params* p, const double* wgt, int L, double tol, int maxit,
                  double* mass, double* temp, double* velo, double* eta) {
  return wto_rk_split(g, p, wgt, L, 4, tol, maxit, mass, temp, velo, eta);
}
