/* CPU oracle of the free-surface (barotropic) elliptic path on one uniform
 * level of the lozenge-tile layout.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * algorithms, used by tests/ and bench.py's CPU baseline as the checker.
 * Nothing in the product links or calls it.
 *
 * Restated routines (reference src/):
 *   elliptic_lo        barotropic_2d.f90:187-253 with step1 itype 2 (ops.f90:202-216, 403-416)
 *                      and div (ops.f90:1325-1343), complete_elliptic_lo (barotropic_2d.f90:240-253)
 *   elliptic_lo_diag   barotropic_2d.f90:255-324 (the "average" branch, exact = .false.)
 *   Jacobi_iteration   lin_solve.f90:271-340
 *   residual, lc       lin_solve.f90:505-618
 *   l2, dp             lin_solve.f90:433-503: sums over the reference's (0,1) sweep, given
 *                      here as a per-cell visit count (red_weight of the grid)
 *   bicgstab           lin_solve.f90:361-431 (van der Vorst 1992)
 *   flux_divergence    barotropic_2d.f90:326-366 with step1 itype 4 (ops.f90:221-235, 422-438)
 *   rhs_elliptic       barotropic_2d.f90:151-185
 * Porosity is 1 (no penalisation), the mean state is zero and every node is
 * active (mask >= ADJZONE).  Evaluation order follows the reference; build
 * with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "elliptic_oracle.h"

static double interp(double a, double b) { return 0.5 * (a + b); }

/* every cell whose value a stencil of an owned cell or pole reads */
static int is_master(const wto_grid* g, int64_t c) {
  int S = g->S, n = g->n;
  int64_t lz = c / ((int64_t)S * S), r = c % ((int64_t)S * S);
  int i = (int)(r % S) - 2, j = (int)(r / S) - 2;
  (void)lz;
  return (i >= 0 && i < n && j >= 0 && j < n) || c == g->south_pole || c == g->north_pole;
}

/* dual and primal lengths of the flux on each edge slot: a halo edge
 * carries its owner's flux (the reference's ghost fill of h_flux) */
static void flux_lengths(const wto_grid* g, double* lenf, double* pedf) {
  const int64_t NC = g->ncells;
  for (int e = 0; e < 3; ++e)
    for (int64_t c = 0; c < NC; ++c) {
      int64_t slot = 3 * c + e, s = g->edge_src[slot];
      int own = !(s >= 0 && s != slot);
      lenf[e * NC + c] = own ? g->len[e * NC + c] : g->len[(s % 3) * NC + s / 3];
      pedf[e * NC + c] = own ? g->pedlen[e * NC + c] : g->pedlen[(s % 3) * NC + s / 3];
    }
}

void wto_node_ghost(const wto_grid* g, double* f) {
  for (int64_t c = 0; c < g->ncells; ++c) {
    int64_t s = g->node_src[c];
    if (s >= 0 && s != c) f[c] = f[s];
  }
}

/* neighbours of cell c in its lozenge frame */
#define NB(c, di, dj) ((c) + (int64_t)(dj) * S + (di))

void wto_elliptic_apply(const wto_grid* g, const wto_fs_params* p, const double* eta_old, const double* q,
                        double* Lq) {
  const int64_t NC = g->ncells;
  const int S = g->S;
  double* lenf = (double*)malloc(3 * NC * sizeof(double));
  double* pedf = (double*)malloc(3 * NC * sizeof(double));
  double* flux = (double*)calloc(3 * NC, sizeof(double));
  flux_lengths(g, lenf, pedf);
  /* csq = g (|topo| phi + eta^n), phi = 1 */
  double* csq = (double*)malloc(NC * sizeof(double));
  for (int64_t c = 0; c < NC; ++c) csq[c] = p->grav * (fabs(g->topo[c]) * 1.0 + eta_old[c]);
  /* fluxes on every edge of a master cell and its W, SW, S neighbours */
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master(g, c)) continue;
    const int64_t cs[4] = {c, NB(c, -1, 0), NB(c, -1, -1), NB(c, 0, -1)};
    for (int a = 0; a < 4; ++a) {
      int64_t x = cs[a];
      int64_t nb[3] = {NB(x, 1, 0), NB(x, 1, 1), NB(x, 0, 1)};
      /* RT: (E - x), DG: (x - NE), UP: (N - x) */
      flux[x] = interp(csq[x], csq[nb[0]]) * (q[nb[0]] / 1.0 - q[x] / 1.0) / lenf[x] * pedf[x];
      flux[NC + x] = interp(csq[x], csq[nb[1]]) * (q[x] / 1.0 - q[nb[1]] / 1.0) / lenf[NC + x] * pedf[NC + x];
      flux[2 * NC + x] = interp(csq[x], csq[nb[2]]) * (q[nb[2]] / 1.0 - q[x] / 1.0) / lenf[2 * NC + x] * pedf[2 * NC + x];
    }
  }
  const double coef = p->theta1 * p->theta2 * (p->dt * p->dt);
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master(g, c)) {
      Lq[c] = 0.0;
      continue;
    }
    int64_t cW = NB(c, -1, 0), cSW = NB(c, -1, -1), cS = NB(c, 0, -1);
    double dv = (flux[c] - flux[cW] + flux[NC + cSW] - flux[NC + c] + flux[2 * NC + c] - flux[2 * NC + cS]) *
                g->hex_inv[c];
    Lq[c] = coef * dv - q[c];
  }
  wto_node_ghost(g, Lq);
  free(lenf);
  free(pedf);
  free(flux);
  free(csq);
}

void wto_elliptic_diag(const wto_grid* g, const wto_fs_params* p, const double* eta_old, double* diag) {
  const int64_t NC = g->ncells;
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master(g, c)) {
      diag[c] = 0.0;
      continue;
    }
    double depth = fabs(g->topo[c]) + eta_old[c] / 1.0;
    double wgt = 2.0 * sqrt(3.0) * depth;
    double lap = -p->grav * wgt * (p->dt * p->dt) * g->hex_inv[c];
    diag[c] = p->theta1 * p->theta2 * lap - 1.0;
  }
  wto_node_ghost(g, diag);
}

void wto_jacobi(const wto_grid* g, const wto_fs_params* p, const double* eta_old, double* u, const double* f,
                double w) {
  const int64_t NC = g->ncells;
  double* Au = (double*)malloc(NC * sizeof(double));
  double* dg = (double*)malloc(NC * sizeof(double));
  wto_elliptic_apply(g, p, eta_old, u, Au);
  wto_elliptic_diag(g, p, eta_old, dg);
  for (int64_t c = 0; c < NC; ++c)
    if (is_master(g, c)) u[c] = u[c] + w * (f[c] - Au[c]) / dg[c];
  wto_node_ghost(g, u);
  free(Au);
  free(dg);
}

double wto_wdot(const wto_grid* g, const double* wgt, const double* a, const double* b) {
  double s = 0.0;
  for (int64_t c = 0; c < g->ncells; ++c)
    if (wgt[c] > 0.0) s += wgt[c] * (a[c] * b[c]);
  return s;
}

static void lc(const wto_grid* g, double* s, double a1, const double* s1, double a2, const double* s2) {
  for (int64_t c = 0; c < g->ncells; ++c)
    if (is_master(g, c)) s[c] = a1 * s1[c] + a2 * s2[c];
  wto_node_ghost(g, s);
}

int wto_bicgstab(const wto_grid* g, const wto_fs_params* p, const double* eta_old, const double* wgt, double* u,
                 const double* f, double nrm_f, double tol, int maxit, double* err_out) {
  const int64_t NC = g->ncells;
  double *res = calloc(NC, 8), *res0 = calloc(NC, 8), *pp = calloc(NC, 8), *s = calloc(NC, 8), *Ap = calloc(NC, 8),
         *As = calloc(NC, 8), *corr = calloc(NC, 8), *Au = calloc(NC, 8);
  int iter;
  double err = 0.0;
  wto_elliptic_apply(g, p, eta_old, u, Au);
  for (int64_t c = 0; c < NC; ++c) res[c] = is_master(g, c) ? f[c] - Au[c] : 0.0;
  wto_node_ghost(g, res);
  memcpy(res0, res, NC * 8);
  memcpy(pp, res0, NC * 8);
  memcpy(s, res0, NC * 8);
  wto_elliptic_apply(g, p, eta_old, pp, Ap);
  memcpy(As, Ap, NC * 8);
  double rho = wto_wdot(g, wgt, res0, res);
  for (iter = 1; iter <= maxit; ++iter) {
    double alph = rho / wto_wdot(g, wgt, Ap, res0);
    lc(g, s, 1.0, res, -alph, Ap);
    wto_elliptic_apply(g, p, eta_old, s, As);
    double omga = wto_wdot(g, wgt, As, s) / wto_wdot(g, wgt, As, As);
    lc(g, corr, alph, pp, omga, s);
    lc(g, u, 1.0, u, 1.0, corr);
    lc(g, res, 1.0, s, -omga, As);
    err = sqrt(wto_wdot(g, wgt, res, res)) / nrm_f;
    if (err <= tol) break;
    double rho_old = rho;
    rho = wto_wdot(g, wgt, res0, res);
    double b = (alph / omga) * (rho / rho_old);
    lc(g, corr, 1.0, pp, -omga, Ap);
    lc(g, pp, 1.0, res, b, corr);
    wto_elliptic_apply(g, p, eta_old, pp, Ap);
  }
  if (err_out) *err_out = err;
  free(res); free(res0); free(pp); free(s); free(Ap); free(As); free(corr); free(Au);
  return iter; /* maxit + 1 when the tolerance was not reached, like the reference's iter_out */
}

void wto_flux_divergence(const wto_grid* g, int L, const double* mass, const double* velo, double* divF) {
  wto_flux_divergence_parts(g, L, mass, velo, divF, NULL);
}

void wto_flux_divergence_parts(const wto_grid* g, int L, const double* mass, const double* velo, double* divF,
                               double* flux_out) {
  const int64_t NC = g->ncells;
  const int S = g->S;
  double* lenf = (double*)malloc(3 * NC * sizeof(double));
  double* pedf = (double*)malloc(3 * NC * sizeof(double));
  double* flux = (double*)calloc(3 * NC, sizeof(double));
  flux_lengths(g, lenf, pedf);
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master(g, c)) continue;
    const int64_t cs[4] = {c, NB(c, -1, 0), NB(c, -1, -1), NB(c, 0, -1)};
    for (int a = 0; a < 4; ++a) {
      int64_t x = cs[a];
      int64_t nb[3] = {NB(x, 1, 0), NB(x, 1, 1), NB(x, 0, 1)};
      for (int e = 0; e < 3; ++e) {
        double h = 0.0;
        for (int k = 0; k < L; ++k) {
          const double* m = mass + (int64_t)k * NC;
          const double* u = velo + (int64_t)k * 3 * NC;
          double dz0 = m[x] + 0.0;
          double dz = interp(dz0, m[nb[e]] + 0.0);
          h = h + u[e * NC + x] * dz;
        }
        flux[e * NC + x] = h * pedf[e * NC + x];
      }
    }
  }
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master(g, c)) {
      divF[c] = 0.0;
      continue;
    }
    int64_t cW = NB(c, -1, 0), cSW = NB(c, -1, -1), cS = NB(c, 0, -1);
    divF[c] = (flux[c] - flux[cW] + flux[NC + cSW] - flux[NC + c] + flux[2 * NC + c] - flux[2 * NC + cS]) *
              g->hex_inv[c];
  }
  wto_node_ghost(g, divF);
  if (flux_out) memcpy(flux_out, flux, 3 * NC * sizeof(double));
  free(lenf);
  free(pedf);
  free(flux);
}

void wto_rhs_elliptic(const wto_grid* g, const wto_fs_params* p, double rho0, const double* eta,
                      const double* divF, const double* divF_old, double* rhs) {
  for (int64_t c = 0; c < g->ncells; ++c)
    rhs[c] = is_master(g, c) ? -eta[c] + p->dt * (p->theta2 * divF[c] + (1.0 - p->theta2) * divF_old[c]) / rho0
                             : 0.0;
  wto_node_ghost(g, rhs);
}
