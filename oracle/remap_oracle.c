/* CPU oracle of the vertical remap of Lagrangian layers (test
 * infrastructure only; see header).
 *
 * Restates, in the evaluation order of the reference's expressions:
 *   remap0 (piecewise constant)                         remap.f90:259-285
 *   remap1 (piecewise linear, harmonic-mean slopes, ten reconciliation sweeps
 *           with a minmod limiter)                      remap.f90:287-380
 *   remap2PPM (Colella-Woodward piecewise parabolic, slope-limited)
 *                                                       remap.f90:382-495
 *   remap_incompressible: old and new interface heights of the node's column
 *     and of the three edge columns (mean of the node's and the neighbour's),
 *     buoyancy and edge velocities remapped, new mass = rho0 dz, new mass-
 *     weighted buoyancy = rho0 dz theta                 remap.f90:147-216
 *   find_coordinates_incompressible (hybrid a/b coordinates)
 *                                                       remap.f90:235-257
 *   apply_no_bdry: owned nodes, then the poles without their velocities
 *                                                       domain.f90:324-341
 */
#include "remap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define H 2
#define CELL(lz, i, j) ((int64_t)(lz) * S * S + (int64_t)((j) + H) * S + (int64_t)((i) + H))

static void remap0(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  double* FC = (double*)malloc(sizeof(double) * (N + 1));
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    FC[k] = fmin(dz, 0.0) * vo[k - 1] + fmax(dz, 0.0) * vo[k];
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = ((zo[k] - zo[k - 1]) * vo[k - 1] + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
  free(FC);
}

/* var_old(k), var_new(k), Hz(k), aL(k), aR(k) of the reference are [k - 1] here */
static void remap1(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  double* FC = (double*)calloc(N + 1, sizeof(double));  /* one layer reads FC(1) before any store */
  double* aL = (double*)malloc(sizeof(double) * N);
  double* aR = (double*)malloc(sizeof(double) * N);
  double* Hz = (double*)malloc(sizeof(double) * N);
  for (int k = 1; k <= N; ++k) Hz[k - 1] = zo[k] - zo[k - 1];
  for (int k = N - 1; k >= 1; --k) FC[k] = (vo[k] - vo[k - 1]) / (Hz[k] + Hz[k - 1]);
  FC[0] = FC[1];
  FC[N] = FC[N - 1];
  for (int k = 1; k <= N; ++k) {
    double cff;
    if (FC[k] * FC[k - 1] < 0.0)
      cff = 0.0;
    else
      cff = 2 * FC[k] * FC[k - 1] / (FC[k] + FC[k - 1]);
    aR[k - 1] = vo[k - 1] + cff * Hz[k - 1];
    aL[k - 1] = vo[k - 1] - cff * Hz[k - 1];
  }
  for (int iter = 1; iter <= 10; ++iter) {
    for (int k = 1; k <= N - 1; ++k) FC[k] = 0.5 * (aR[k - 1] + aL[k]);
    FC[N] = aR[N - 1];
    FC[0] = aL[0];
    for (int k = 1; k <= N; ++k) {
      const double dR = FC[k] - vo[k - 1], dL = vo[k - 1] - FC[k - 1];
      double cff;
      if (dR * dL < 0.0)
        cff = 0.0;
      else if (fabs(dR) > fabs(dL))
        cff = dL;
      else
        cff = dR;
      aR[k - 1] = vo[k - 1] + cff;
      aL[k - 1] = vo[k - 1] - cff;
    }
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    double cff, cff1, dh;
    if (dz > 0.0) {
      cff = aL[k];
      cff1 = aR[k] - aL[k];
      dh = Hz[k];
    } else {
      cff = aR[k - 1];
      cff1 = aR[k - 1] - aL[k - 1];
      dh = Hz[k - 1];
    }
    FC[k] = dz * (cff + 0.5 * cff1 * dz / dh);
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k - 1] * vo[k - 1] + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
  free(FC);
  free(aL);
  free(aR);
  free(Hz);
}

/* remap2PPM (remap.f90:382-495): arrays (0:N) of the reference at [k];
 * var(k), Hz(k) at [k - 1].  FC holds slopes, then curvatures, then (in
 * increasing k, as the reference overwrites it in place) the fluxes. */
static void remap2ppm(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  double *Hz = calloc(N + 1, sizeof(double)), *aL = calloc(N + 1, sizeof(double)), *aR = calloc(N + 1, sizeof(double));
  double *CF = calloc(N + 1, sizeof(double)), *FC = calloc(N + 1, sizeof(double)), *FC1 = calloc(N + 1, sizeof(double));
#define V(k) vo[(k) - 1]
#define HZ(k) Hz[(k) - 1]
  for (int k = 1; k <= N; ++k) HZ(k) = zo[k] - zo[k - 1];
  for (int k = 1; k <= N - 1; ++k) {
    const double cff = 1.0 / (HZ(k) + HZ(k + 1));
    CF[k] = cff * (V(k + 1) * HZ(k) + V(k) * HZ(k + 1));
    FC[k] = cff * (V(k + 1) - V(k));
  }
  for (int k = 2; k <= N - 1; ++k) {
    double cff = HZ(k) * ((2.0 * HZ(k - 1) + HZ(k)) * FC[k] + (2.0 * HZ(k + 1) + HZ(k)) * FC[k - 1]) /
                 (HZ(k - 1) + HZ(k) + HZ(k + 1));
    const double cffR = 2.0 * (V(k + 1) - V(k)), cffL = 2.0 * (V(k) - V(k - 1));
    if (cffR * cffL > 0.0) {
      if (fabs(cffL) < fabs(cff)) cff = cffL;
      if (fabs(cffR) < fabs(cff)) cff = cffR;
    } else {
      cff = 0.0;
    }
    FC1[k] = cff;
  }
  FC1[N] = HZ(N) * FC[N - 1];
  FC1[1] = HZ(1) * FC[1];
  for (int k = 1; k <= N - 1; ++k) {
    const int k1 = k - 1 > 1 ? k - 1 : 1, k2 = k + 2 < N ? k + 2 : N;
    const double cff = HZ(k1) + HZ(k) + HZ(k + 1) + HZ(k2);
    const double cffL = (HZ(k1) + HZ(k)) / (cff * (2.0 * HZ(k) + HZ(k + 1)));
    const double cffR = (HZ(k + 1) + HZ(k2)) / (cff * (HZ(k) + 2.0 * HZ(k + 1)));
    aR[k] = CF[k] + 2.0 * HZ(k) * HZ(k + 1) * (cffL - cffR) * FC[k] - cffL * HZ(k) * FC1[k + 1] +
            cffR * HZ(k + 1) * FC1[k];
    aL[k + 1] = aR[k];
  }
  aR[N] = 2.0 * V(N) - aL[N];
  aL[1] = 2.0 * V(1) - aR[1];
  for (int k = 1; k <= N; ++k) {
    CF[k] = 0.5 * (aR[k] - aL[k]);
    FC[k] = 0.5 * (aR[k] + aL[k]) - V(k);
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    double alpha, cff, cffR, cffL;
    if (dz > 0.0) {
      alpha = HZ(k + 1);
      cff = aL[k + 1];
      cffR = CF[k + 1];
      cffL = FC[k + 1];
    } else {
      alpha = -HZ(k);
      cff = aR[k];
      cffR = -CF[k];
      cffL = FC[k];
    }
    alpha = dz / alpha;
    FC[k] = dz * (cff + alpha * (cffR - cffL * (3.0 - 2.0 * alpha)));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (HZ(k) * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
#undef V
#undef HZ
  free(Hz); free(aL); free(aR); free(CF); free(FC); free(FC1);
}

void wto_remap_column(int scheme, int N, double* vn, const double* zn, const double* vo, const double* zo) {
  if (scheme == 0)
    remap0(N, vn, zn, vo, zo);
  else if (scheme == 1)
    remap1(N, vn, zn, vo, zo);
  else
    remap2ppm(N, vn, zn, vo, zo);
}

/* find_coordinates_incompressible of column c from the saved masses */
static void coordinates(const wto_grid* g, int L, const double* a_vert, const double* b_vert, const double* m_old,
                        int64_t c, double* zn, double* zo) {
  const int64_t NC = g->ncells;
  const double zs = g->topo[c];
  zo[0] = zs;
  for (int k = 1; k <= L; ++k) {
    const double rho_dz = 0.0 + m_old[(int64_t)(k - 1) * NC + c];
    zo[k] = zo[k - 1] + rho_dz / (g->rho0 * 1.0);
  }
  const double eta = zo[L];
  for (int k = 0; k <= L; ++k) zn[k] = a_vert[k] * eta + b_vert[k] * zs;
}

static void remap_column_node(const wto_grid* g, int scheme, int L, const double* a_vert, const double* b_vert,
                              const double* m_old, int64_t c, int pole, double* mass, double* temp, double* velo) {
  const int S = g->S;
  const int64_t NC = g->ncells;
  double *zn = malloc(sizeof(double) * (L + 1)), *zo = malloc(sizeof(double) * (L + 1));
  double *zen = malloc(sizeof(double) * (L + 1)), *zeo = malloc(sizeof(double) * (L + 1));
  double *th_o = malloc(sizeof(double) * L), *th_n = malloc(sizeof(double) * L);
  double *fo = malloc(sizeof(double) * 3 * L), *fn = malloc(sizeof(double) * 3 * L);
  double* ezn = malloc(sizeof(double) * 3 * (L + 1));
  double* ezo = malloc(sizeof(double) * 3 * (L + 1));
  coordinates(g, L, a_vert, b_vert, m_old, c, zn, zo);
  const int64_t nb[3] = {c + 1, c + S + 1, c + S};
  for (int e = 0; e < 3; ++e) {
    coordinates(g, L, a_vert, b_vert, m_old, nb[e], zen, zeo);
    for (int k = 0; k <= L; ++k) {
      ezn[e * (L + 1) + k] = 0.5 * (zn[k] + zen[k]);
      ezo[e * (L + 1) + k] = 0.5 * (zo[k] + zeo[k]);
    }
  }
  for (int k = 0; k < L; ++k) {
    const int64_t x = (int64_t)k * NC + c;
    th_o[k] = (0.0 + temp[x]) / (0.0 + mass[x]);
    for (int e = 0; e < 3; ++e) fo[e * L + k] = velo[((int64_t)k * 3 + e) * NC + c];
  }
  wto_remap_column(scheme, L, th_n, zn, th_o, zo);
  for (int e = 0; e < 3; ++e)
    wto_remap_column(scheme, L, fn + e * L, ezn + e * (L + 1), fo + e * L, ezo + e * (L + 1));
  for (int k = 0; k < L; ++k) {
    const int64_t x = (int64_t)k * NC + c;
    const double rho = g->rho0 * 1.0;
    const double rho_dz = rho * (zn[k + 1] - zn[k]);
    mass[x] = rho_dz - 0.0;
    temp[x] = rho_dz * th_n[k] - 0.0;
    if (!pole)
      for (int e = 0; e < 3; ++e) velo[((int64_t)k * 3 + e) * NC + c] = fn[e * L + k];
  }
  free(zn); free(zo); free(zen); free(zeo); free(th_o); free(th_n); free(fo); free(fn); free(ezn); free(ezo);
}

void wto_remap(const wto_grid* g, int scheme, int L, const double* a_vert, const double* b_vert, double* mass,
               double* temp, double* velo) {
  const int n = g->n, S = g->S;
  const int64_t NC = g->ncells;
  double* m_old = (double*)malloc(sizeof(double) * (size_t)L * NC);
  memcpy(m_old, mass, sizeof(double) * (size_t)L * NC);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        remap_column_node(g, scheme, L, a_vert, b_vert, m_old, CELL(lz, i, j), 0, mass, temp, velo);
  remap_column_node(g, scheme, L, a_vert, b_vert, m_old, g->south_pole, 1, mass, temp, velo);
  remap_column_node(g, scheme, L, a_vert, b_vert, m_old, g->north_pole, 1, mass, temp, velo);
  free(m_old);
  wto_ghost_fill(g, L, mass, temp, velo);
}
