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
 *   remap2S (parabolic spline, parabolic continuation)  remap.f90:497-583
 *   remap2W (parabolic WENO, limited interior)          remap.f90:585-735
 *   remap4 (parabolic WENO with power-law reconciliation: quartic profiles)
 *                                                       remap.f90:737-948
 *   remapPPR (Engwirda's PPR library at order 2: p3e cubic edge values with
 *           loose end fits, PPM cells with the mono limiter, integral
 *           remap with the compensated conservation fix)
 *                                                       remap.f90:950-1010,
 *           ppr/rmap1d.f90, rcon1d.f90, p3e.f90, p1e.f90, pbc.f90,
 *           ppm.f90, plm.f90 (plsv, plsc), inv.f90 (slv_3x3, slv_4x4)
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

/* The three schemes below use the reference's (1:N) and (0:N) arrays at the
 * same indices: V(k) is var_old(k); Hz, FC, r, ... hold N + 2 entries (zeroed:
 * a one-layer column reads entries the reference never sets, which do not
 * reach its result). */
#define V(k) vo[(k) - 1]

/* remap2S (remap.f90:497-583): parabolic spline with parabolic continuation
 * at both ends; at least two layers (one layer is the flux-free update) */
static void remap2s(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  double* w = (double*)calloc(6 * (size_t)(N + 2), sizeof(double));
  double *Hz = w, *dL = w + (N + 2), *dR = w + 2 * (N + 2), *FC = w + 3 * (N + 2), *r = w + 4 * (N + 2);
  for (int k = 1; k <= N; ++k) Hz[k] = zo[k] - zo[k - 1];
  if (N >= 2) {
    double cff = Hz[1] / Hz[2];
    FC[1] = 1.0 + cff;
    r[0] = 2.0 * V(1) + cff * (V(1) + cff * V(2)) / FC[1];
    for (int k = 1; k <= N - 1; ++k) {
      cff = 1.0 / (2.0 * Hz[k] + Hz[k + 1] * (2.0 - FC[k]));
      FC[k + 1] = cff * Hz[k];
      r[k] = cff * (3.0 * (V(k + 1) * Hz[k] + V(k) * Hz[k + 1]) - Hz[k + 1] * r[k - 1]);
    }
    cff = Hz[N] / Hz[N - 1];
    const double cff1 = 1.0 + cff;
    r[N] = (cff * (V(N) + cff * V(N - 1)) + cff1 * (2.0 * V(N) - cff1 * r[N - 1])) / (cff1 * (1.0 - cff1 * FC[N]));
    for (int k = N - 1; k >= 0; --k) r[k] = r[k] - FC[k + 1] * r[k + 1];
  }
  for (int k = 1; k <= N; ++k) {
    dR[k] = 0.5 * (r[k] - r[k - 1]);
    dL[k] = 0.5 * (r[k] + r[k - 1]) - V(k);
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    double alpha, cff, cff1;
    if (dz > 0.0) {
      alpha = Hz[k + 1]; cff = dR[k + 1]; cff1 = dL[k + 1];
    } else {
      alpha = Hz[k]; cff = dR[k]; cff1 = -dL[k];
    }
    alpha = dz / alpha;
    FC[k] = dz * (r[k] + alpha * (cff - cff1 * (3.0 - 2.0 * fabs(alpha))));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k] * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
  free(w);
}

/* parabolic WENO side limits shared by remap2W and remap4 (remap.f90:605-639,
 * 760-795): monotonised parabolic aL, aR and their quadratic variations dL,
 * dR in the inner cells; d holds the centred differences (remap4's FC * dv,
 * remap2W's FC) */
static void weno_sides(int N, const double* vo, const double* Hz, const double* d, double* aL, double* aR, double* dL,
                       double* dR) {
  for (int k = 2; k <= N - 1; ++k) {
    double deltaR = Hz[k] * d[k], deltaL = Hz[k] * d[k - 1];
    if (deltaR * deltaL < 0.0) {
      deltaR = 0.0;
      deltaL = 0.0;
    }
    double cff = Hz[k - 1] + 2.0 * Hz[k] + Hz[k + 1];
    const double cffR = cff * d[k], cffL = cff * d[k - 1];
    if (fabs(deltaR) > fabs(cffL)) deltaR = cffL;
    if (fabs(deltaL) > fabs(cffR)) deltaL = cffR;
    cff = (deltaR - deltaL) / (Hz[k - 1] + Hz[k] + Hz[k + 1]);
    deltaR = deltaR - cff * Hz[k + 1];
    deltaL = deltaL + cff * Hz[k - 1];
    aR[k] = V(k) + deltaR;
    aL[k] = V(k) - deltaL;
    const double pr = 2.0 * deltaR - deltaL, pl = 2.0 * deltaL - deltaR;
    dR[k] = pr * pr;
    dL[k] = pl * pl;
  }
}

/* remap2W (remap.f90:585-735) with LIMIT_INTERIOR and no NEUMANN */
static void remap2w(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  double* w = (double*)calloc(7 * (size_t)(N + 2), sizeof(double));
  double *Hz = w, *aL = w + (N + 2), *aR = w + 2 * (N + 2), *dL = w + 3 * (N + 2), *dR = w + 4 * (N + 2);
  double *FC = w + 5 * (N + 2), *r = w + 6 * (N + 2);
  const double eps = 1e-8;
  for (int k = 1; k <= N; ++k) Hz[k] = zo[k] - zo[k - 1];
  for (int k = N - 1; k >= 1; --k) FC[k] = (V(k + 1) - V(k)) / (Hz[k + 1] + Hz[k]);
  weno_sides(N, vo, Hz, FC, aL, aR, dL, dR);
  aR[N] = V(N); aL[N] = V(N); dR[N] = 0.0; dL[N] = 0.0;
  aR[1] = V(1); aL[1] = V(1); dR[1] = 0.0; dL[1] = 0.0;
  for (int k = 1; k <= N - 1; ++k) {
    const double deltaL = fmax(dL[k], eps), deltaR = fmax(dR[k + 1], eps);
    r[k] = (deltaR * aR[k] + deltaL * aL[k + 1]) / (deltaR + deltaL);
  }
  r[N] = 2.0 * V(N) - r[N - 1];
  r[0] = 2.0 * V(1) - r[1];
  for (int k = 1; k <= N; ++k) {
    double deltaR = r[k] - V(k), deltaL = V(k) - r[k - 1];
    const double cffR = 2.0 * deltaR, cffL = 2.0 * deltaL;
    if (deltaR * deltaL < 0.0) {
      deltaR = 0.0;
      deltaL = 0.0;
    } else if (fabs(deltaR) > fabs(cffL)) {
      deltaR = cffL;
    } else if (fabs(deltaL) > fabs(cffR)) {
      deltaL = cffR;
    }
    aR[k] = V(k) + deltaR;
    aL[k] = V(k) - deltaL;
    dL[k] = 0.5 * (aR[k] - aL[k]);
    dR[k] = 0.5 * (aR[k] + aL[k]) - V(k);
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    double alpha, cff, cffL, cffR;
    if (dz > 0.0) {
      alpha = Hz[k + 1]; cff = aL[k + 1]; cffL = dL[k + 1]; cffR = dR[k + 1];
    } else {
      alpha = -Hz[k]; cff = aR[k]; cffL = -dL[k]; cffR = dR[k];
    }
    alpha = dz / alpha;
    FC[k] = dz * (cff + alpha * (cffL - cffR * (3.0 - 2.0 * alpha)));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k] * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
  free(w);
}

/* remap4 (remap.f90:737-948): parabolic WENO, then the power-law
 * reconciliation of interface values and derivatives (quartic profiles),
 * without NEUMANN; at least three layers (two read an unset value in the
 * reference).  Its Pade constants 0.0763..., 0.004329... and 0.0363... are
 * default-kind (single precision) literals in the reference. */
static void remap4(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  double* w = (double*)calloc(10 * (size_t)(N + 2), sizeof(double));
  double *Hz = w, *aL = w + (N + 2), *aR = w + 2 * (N + 2), *dL = w + 3 * (N + 2), *dR = w + 4 * (N + 2);
  double *d = w + 5 * (N + 2), *FC = w + 6 * (N + 2), *r = w + 7 * (N + 2), *r1 = w + 8 * (N + 2);
  const double eps = 1e-8;
  const double c1 = (double)0.0763636363636363636f, c2 = (double)0.004329004329004329f;
  const double c3 = (double)0.0363636363636363636f;
  for (int k = 1; k <= N; ++k) Hz[k] = zo[k] - zo[k - 1];
  for (int k = N - 1; k >= 1; --k) {
    FC[k] = 1.0 / (Hz[k + 1] + Hz[k]);
    d[k] = FC[k] * (V(k + 1) - V(k));
  }
  weno_sides(N, vo, Hz, d, aL, aR, dL, dR);
  {
    aL[N] = aR[N - 1];
    aR[N] = 2.0 * V(N) - aL[N];
    const double pr = 2.0 * aR[N] + aL[N] - 3.0 * V(N), pl = 3.0 * V(N) - 2.0 * aL[N] - aR[N];
    dR[N] = pr * pr;
    dL[N] = pl * pl;
  }
  if (N >= 2) {
    aR[1] = aL[2];
    aL[1] = 2.0 * V(1) - aR[1];
    const double pr = 2.0 * aR[1] + aL[1] - 3.0 * V(1), pl = 3.0 * V(1) - 2.0 * aL[1] - aR[1];
    dR[1] = pr * pr;
    dL[1] = pl * pl;
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double deltaL = fmax(dL[k], eps), deltaR = fmax(dR[k + 1], eps);
    r1[k] = (deltaR * aR[k] + deltaL * aL[k + 1]) / (deltaR + deltaL);
  }
  r1[N] = 2.0 * V(N) - r1[N - 1];
  r1[0] = 2.0 * V(1) - r1[1];
  for (int k = 1; k <= N; ++k) {
    const double deltaR = r1[k] - V(k), deltaL = V(k) - r1[k - 1];
    double cff = deltaR * deltaL;
    cff = cff > eps ? (deltaR + deltaL) / cff : 0.0;
    double cffL = cff * deltaL, cffR = cff * deltaR;
    if (cffL > 3.0) {
      cffL = cffL * deltaL;
      cffR = 0.0;
    } else if (cffR > 3.0) {
      cffL = 0.0;
      cffR = cffR * deltaR;
    } else {
      cffL = 4.0 * deltaL - 2.0 * deltaR;
      cffR = 4.0 * deltaR - 2.0 * deltaL;
    }
    cff = 1.0 / Hz[k];
    dR[k] = cff * cffR;
    dL[k] = cff * cffL;
  }
  for (int k = N - 1; k >= 1; --k) {
    d[k] = FC[k] * (Hz[k + 1] * dL[k + 1] + Hz[k] * dR[k]);
    const double cffR = 8.0 * (dR[k] + 2.0 * dL[k]), cffL = 8.0 * (dL[k + 1] + 2.0 * dR[k + 1]);
    if (fabs(d[k]) > fabs(cffR)) d[k] = cffR;
    if (fabs(d[k]) > fabs(cffL)) d[k] = cffL;
    double Hdd, rr;
    if ((dL[k + 1] - dR[k]) * (V(k + 1) - V(k)) > 0.0) {
      Hdd = Hz[k] * (d[k] - dR[k]);
      rr = V(k) - r1[k - 1];
    } else {
      Hdd = Hz[k + 1] * (dL[k + 1] - d[k]);
      rr = r1[k + 1] - V(k + 1);
    }
    rr = fabs(rr);
    double Ampl = 0.2 * Hdd * rr;
    Hdd = fabs(Hdd);
    const double cff = rr * rr + c1 * Hdd * (rr + c2 * Hdd);
    Ampl = cff > eps ? Ampl * (rr + c3 * Hdd) / cff : 0.0;
    r[k] = r1[k] + Ampl;
  }
  r[0] = 2.0 * V(1) - r[1];
  r[N] = 2.0 * V(N) - r[N - 1];
  d[0] = d[1];
  d[N] = d[N - 1];
  for (int k = 1; k <= N; ++k) {
    aR[k] = r[k] + r[k - 1] - 2.0 * V(k);
    dR[k] = Hz[k] * (d[k] + d[k - 1]) - 2.0 * (r[k] - r[k - 1]);
    dL[k] = Hz[k] * (d[k] - d[k - 1]) - 6.0 * aR[k];
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    double alpha, cff, cffR, cffL;
    if (dz > 0.0) {
      alpha = Hz[k + 1]; cff = aR[k + 1]; cffR = dR[k + 1]; cffL = dL[k + 1];
    } else {
      alpha = -Hz[k]; cff = aR[k]; cffR = -dR[k]; cffL = dL[k];
    }
    alpha = dz / alpha;
    FC[k] = dz * (r[k] + 0.5 * dz * d[k] +
                  alpha * alpha * (cff - cffR * (0.5 - 0.25 * alpha) + cffL * (1.0 - alpha * (1.25 - 0.5 * alpha))));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k] * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
  free(w);
}
#undef V

/* ---- remapPPR (remap.f90:950-1010): Engwirda's PPR library at order 2 ----
 * cubic edge values (p3e.f90) with loose one-sided end fits (pbc.f90 lbc/ubc),
 * monotone piecewise-parabolic cells (ppm.f90 ppm/ppmfn, plm.f90 plsv/plsc),
 * exact integration of the cell polynomials over the new cells with the
 * compensated conservation fix (rmap1d.f90 rmap1d/imap1d).  Arrays below are
 * indexed as the library's: edges 1..N+1, cells 1..N (element 0 unused). */

/* integer powers as the reference's compiled x**k evaluates them: products
 * from the left (x**4 as a square of squares differs in the last bit) */
static double ppr_p2(double x) { return x * x; }
static double ppr_p3(double x) { return (x * x) * x; }
static double ppr_p4(double x) { return ((x * x) * x) * x; }

/* slv_3x3 (inv.f90:64-124, 400-445): adjugate and determinant, solve in place */
static int ppr_solve3(const double A[3][3], double v[3], double eps) {
  const double a2233 = A[1][1] * A[2][2], a2332 = A[1][2] * A[2][1], a2133 = A[1][0] * A[2][2];
  const double a2331 = A[1][2] * A[2][0], a2132 = A[1][0] * A[2][1], a2231 = A[1][1] * A[2][0];
  const double det = A[0][0] * (a2233 - a2332) - A[0][1] * (a2133 - a2331) + A[0][2] * (a2132 - a2231);
  const double a1233 = A[0][1] * A[2][2], a1332 = A[0][2] * A[2][1], a1223 = A[0][1] * A[1][2];
  const double a1322 = A[0][2] * A[1][1], a1133 = A[0][0] * A[2][2], a1331 = A[0][2] * A[2][0];
  const double a1123 = A[0][0] * A[1][2], a1321 = A[0][2] * A[1][0], a1132 = A[0][0] * A[2][1];
  const double a1231 = A[0][1] * A[2][0], a1122 = A[0][0] * A[1][1], a1221 = A[0][1] * A[1][0];
  const double adj[3][3] = {{(a2233 - a2332), -(a1233 - a1332), (a1223 - a1322)},
                            {-(a2133 - a2331), (a1133 - a1331), -(a1123 - a1321)},
                            {(a2132 - a2231), -(a1132 - a1231), (a1122 - a1221)}};
  if (!(fabs(det) > eps)) return 0;
  double t[3];
  for (int r = 0; r < 3; ++r) t[r] = (adj[r][0] * v[0] + adj[r][1] * v[1] + adj[r][2] * v[2]) / det;
  v[0] = t[0]; v[1] = t[1]; v[2] = t[2];
  return 1;
}

/* C += s (A B) for 2 x 2 blocks (inv.f90 mul_2x2; its +1 / -1 branches are the
 * same numbers as the general one) */
static void ppr_mul2(const double A[2][2], const double B[2][2], double s, double C[2][2]) {
  for (int c = 0; c < 2; ++c)
    for (int r = 0; r < 2; ++r) {
      const double p = A[r][0] * B[0][c] + A[r][1] * B[1][c];
      C[r][c] = s == 1.0 ? C[r][c] + p : s == -1.0 ? C[r][c] - p : C[r][c] + s * p;
    }
}

/* slv_4x4 (inv.f90:447-600): block LDU through 2 x 2 adjugates; the second
 * singularity test repeats the first (as the library does) */
static int ppr_solve4(const double A[4][4], double v[4], double eps) {
  const double A11[2][2] = {{A[0][0], A[0][1]}, {A[1][0], A[1][1]}}, A12[2][2] = {{A[0][2], A[0][3]}, {A[1][2], A[1][3]}};
  const double A21[2][2] = {{A[2][0], A[2][1]}, {A[3][0], A[3][1]}};
  const double adet = A11[0][0] * A11[1][1] - A11[0][1] * A11[1][0];
  const double ai[2][2] = {{A11[1][1], -A11[0][1]}, {-A11[1][0], A11[0][0]}};
  if (!(fabs(adet) > eps)) return 0;
  double L[2][2] = {{0, 0}, {0, 0}}, U[2][2] = {{0, 0}, {0, 0}};
  double Sm[2][2] = {{A[2][2], A[2][3]}, {A[3][2], A[3][3]}};
  ppr_mul2(A21, ai, 1.0, L);
  ppr_mul2(ai, A12, 1.0, U);
  ppr_mul2(L, A12, -1.0 / adet, Sm);
  const double sdet = Sm[0][0] * Sm[1][1] - Sm[0][1] * Sm[1][0];
  const double si[2][2] = {{Sm[1][1], -Sm[0][1]}, {-Sm[1][0], Sm[0][0]}};
  v[2] = v[2] - (L[0][0] * v[0] + L[0][1] * v[1]) / adet;
  v[3] = v[3] - (L[1][0] * v[0] + L[1][1] * v[1]) / adet;
  const double t0 = (ai[0][0] * v[0] + ai[0][1] * v[1]) / adet, t1 = (ai[1][0] * v[0] + ai[1][1] * v[1]) / adet;
  v[0] = t0; v[1] = t1;
  const double t2 = (si[0][0] * v[2] + si[0][1] * v[3]) / sdet, t3 = (si[1][0] * v[2] + si[1][1] * v[3]) / sdet;
  v[2] = t2; v[3] = t3;
  v[0] = v[0] - (U[0][0] * v[2] + U[0][1] * v[3]) / adet;
  v[1] = v[1] - (U[1][0] * v[2] + U[1][1] * v[3]) / adet;
  return 1;
}

/* loose end fit (pbc.f90 lbc / ubc): the quadratic whose means match the three
 * cells c-1, c, c+1, evaluated at the edges c-1 .. c+1 (low end, lo = 1) or
 * c .. c+2 (high end, lo = 0); dx has nd = 1 entry when the column is uniform */
static void ppr_end_fit(int lo, int c, const double* dx, int nd, const double* f, double dmin, double* edge) {
  double xhat, h[3];
  if (nd > 1) {
    xhat = fmax(dx[c], dmin) * 0.5;
    h[0] = dx[c - 1]; h[1] = dx[c]; h[2] = dx[c + 1];
  } else {
    xhat = fmax(dx[1], dmin) * 0.5;
    h[0] = h[1] = h[2] = dx[1];
  }
  const double s[4] = {-(h[0] + h[1] * 0.5) / xhat, -1.0, 1.0, (h[2] + h[1] * 0.5) / xhat};
  double A[3][3], v[3];
  for (int r = 0; r < 3; ++r) {
    A[r][0] = s[r + 1] / 1.0 - s[r] / 1.0;
    A[r][1] = ppr_p2(s[r + 1]) / 2.0 - ppr_p2(s[r]) / 2.0;
    A[r][2] = ppr_p3(s[r + 1]) / 3.0 - ppr_p3(s[r]) / 3.0;
    v[r] = h[r] * f[c - 1 + r] / xhat;
  }
  const int e0 = lo ? c - 1 : c;  /* first edge written */
  if (ppr_solve3(A, v, 1.0e-14 * dmin)) {
    for (int k = 0; k < 3; ++k) {
      const double x = s[k + (lo ? 0 : 1)];
      edge[e0 + k] = 1.0 * v[0] + (x * 1.0) * v[1] + (ppr_p2(x) * 1.0) * v[2];
    }
  } else if (lo) {
    edge[c - 1] = f[c - 1] * 1.0;
    edge[c] = f[c - 1] * 0.5 + f[c] * 0.5;
    edge[c + 1] = f[c] * 0.5 + f[c + 1] * 0.5;
  } else {
    edge[c] = f[c - 1] * 0.5 + f[c] * 0.5;
    edge[c + 1] = f[c] * 0.5 + f[c + 1] * 0.5;
    edge[c + 2] = f[c + 1] * 1.0;
  }
}

/* edge values: p3e.f90 (p1e.f90 for at most three cells) */
static void ppr_edges(int N, const double* dx, int nd, const double* f, double dmin, double* edge) {
  const int npos = N + 1;
  if (npos <= 4) {  /* p1e: two-point means inside, the end cells' values at the ends */
    for (int i = 2; i <= npos - 1; ++i) {
      if (nd == 1) {
        const double d = dx[1] * 2.0;
        edge[i] = (dx[1] * f[i - 1] + dx[1] * f[i]) / d;
      } else {
        const double hl = fmax(dx[i - 1], dmin), h0 = fmax(dx[i], dmin);
        edge[i] = (h0 * f[i - 1] + hl * f[i]) / (hl + h0);
      }
    }
    edge[1] = f[1];
    edge[npos] = f[npos - 1];
    return;
  }
  ppr_end_fit(1, 2, dx, nd, f, dmin, edge);
  ppr_end_fit(0, npos - 2, dx, nd, f, dmin, edge);
  for (int i = 3; i <= npos - 2; ++i) {
    if (nd == 1) {
      edge[i] = (-(1.0 * f[i - 2]) + 7.0 * f[i - 1] + 7.0 * f[i] - 1.0 * f[i + 1]) / 12.0;
      continue;
    }
    const double h[4] = {dx[i - 2], dx[i - 1], dx[i], dx[i + 1]};
    const double xhat = 0.5 * fmax(h[1], dmin) + 0.5 * fmax(h[2], dmin);
    const double s[5] = {-(h[0] + h[1]) / xhat, -h[1] / xhat, 0.0, h[2] / xhat, (h[2] + h[3]) / xhat};
    double iv[5][4];
    for (int k = 0; k < 5; ++k) {
      iv[k][0] = s[k] / 1.0;
      iv[k][1] = ppr_p2(s[k]) / 2.0;
      iv[k][2] = ppr_p3(s[k]) / 3.0;
      iv[k][3] = ppr_p4(s[k]) / 4.0;
    }
    double A[4][4], v[4];
    for (int r = 0; r < 4; ++r) {
      for (int p = 0; p < 4; ++p) A[r][p] = iv[r + 1][p] - iv[r][p];
      v[r] = dx[i - 2 + r] * f[i - 2 + r] / xhat;
    }
    if (ppr_solve4(A, v, 1.0e-14 * dmin))
      edge[i] = v[0];
    else
      edge[i] = f[i - 1] * 0.5 + f[i] * 0.5;
  }
}

/* limited slope of the cell (plm.f90 plsv for a non-uniform column, plsc for a
 * uniform one): only the centred slope dfds(0) is used by ppmfn */
static double ppr_slope(int uniform, double fl, double hl, double f0, double h0, double fr, double hr) {
  const double dm = f0 - fl, dp = fr - f0;
  if (!(dm * dp > 0.0)) return 0.0;
  double el, er;
  if (uniform) {
    el = (fl + f0) * 0.5;
    er = (f0 + fr) * 0.5;
  } else {
    el = (h0 * fl + hl * f0) / (hl + h0);
    er = (hr * f0 + h0 * fr) / (h0 + hr);
  }
  const double d0 = 0.5 * (er - el);
  double scal = fmin(fabs(dm), fabs(dp)) / fmax(fabs(d0), 1.0e-14);
  scal = fmin(scal, 1.0);
  return scal * d0;
}

static void ppr_parabola(double f0, double el, double er, double c[3]) {
  c[0] = 1.5 * f0 - 0.25 * (er + el);
  c[1] = 0.5 * (er - el);
  c[2] = -(1.5 * f0) + 0.75 * (er + el);
}

/* ppm.f90 ppm + ppmfn with the mono limiter: coefficients of 1, s, s^2 on the
 * cell mapped to s in [-1, 1] */
static void ppr_cells(int N, const double* dx, int nd, const double* f, const double* edge, double (*fh)[3]) {
  if (N == 1) {
    fh[1][0] = f[1]; fh[1][1] = 0.0; fh[1][2] = 0.0;
    return;
  }
  for (int i = 1; i <= N; ++i) {
    const int il = i - 1 > 1 ? i - 1 : 1, ir = i + 1 < N ? i + 1 : N;
    const double f0 = f[i], fl = f[il], fr = f[ir];
    double el = edge[i], er = edge[i + 1];
    const double d0 = nd > 1 ? ppr_slope(0, fl, dx[il], f0, dx[i], fr, dx[ir]) : ppr_slope(1, fl, 0, f0, 0, fr, 0);
    double u[3], l[3];
    ppr_parabola(f0, el, er, u);
    if ((fr - f0) * (f0 - fl) < 0.0) {
      l[0] = f0; l[1] = 0.0; l[2] = 0.0;
    } else {
      if ((fl - el) * (el - f0) <= 0.0) el = f0 - d0;
      if ((fr - er) * (er - f0) <= 0.0) er = f0 + d0;
      ppr_parabola(f0, el, er, l);
      if (fabs(l[2]) > fabs(l[1]) * 0.5) {
        const double turn = -0.5 * l[1] / l[2];
        if (turn >= -1.0 && turn <= 0.0) {
          er = 3.0 * f0 - 2.0 * el;
          ppr_parabola(f0, el, er, l);
        } else if (turn > 0.0 && turn <= 1.0) {
          el = 3.0 * f0 - 2.0 * er;
          ppr_parabola(f0, el, er, l);
        }
      }
    }
    for (int p = 0; p < 3; ++p) fh[i][p] = 0.0 * u[p] + 1.0 * l[p];
  }
}

/* add b to the compensated sum (s, e) (imap1d's Kahan-Babuska-Neumaier loop) */
static void ppr_kbn(double* s, double* e, double b) {
  const double t = *s + b;
  if (fabs(*s) >= fabs(b))
    *e = *e + ((*s - t) + b);
  else
    *e = *e + ((b - t) + *s);
  *s = t;
}

static void remap_ppr(int N, double* vn, const double* zn, const double* vo, const double* zo) {
  const int npos = N + 1;
  const double* x = zo - 1;   /* x[1..npos] */
  const double* xn = zn - 1;
  const double* f = vo - 1;   /* f[1..N] */
  double* dx = (double*)calloc(npos + 1, sizeof(double));
  double* edge = (double*)calloc(npos + 1, sizeof(double));
  double (*fh)[3] = (double (*)[3])calloc(npos + 1, sizeof(double[3]));
  double* fnew = (double*)calloc(npos + 1, sizeof(double));
  /* rmap1d: uniform spacing to 1e-14 keeps one spacing */
  const double same = (x[npos] - x[1]) / (npos - 1), xtol = same * 1.0e-14;
  int uniform = 1;
  for (int i = 1; i <= npos - 1; ++i) {
    dx[i] = x[i + 1] - x[i];
    if (fabs(dx[i] - same) > xtol) uniform = 0;
  }
  const int nd = uniform ? 1 : N;
  /* rcon1d: dmin = 1e-14 times the mean spacing */
  double dmid = dx[1];
  if (nd > 1) {
    for (int i = 2; i <= npos - 1; ++i) dmid = dmid + dx[i];
    dmid = dmid / (npos - 1);
  }
  const double dmin = 1.0e-14 * dmid;
  if (N >= 2) ppr_edges(N, dx, nd, f, dmin, edge);
  ppr_cells(N, dx, nd, f, edge, fh);
  /* imap1d with three degrees of freedom */
  int kmin = 1, kmax = 1, p0 = 1, p1 = 1;
  for (int k = 1; k <= N; ++k) {
    if (p1 < 1) p1 = 1;
    for (p0 = p1; p0 <= npos - 1; ++p0)
      if (x[p0 + 1] > xn[k]) break;
    for (p1 = p0; p1 <= npos - 1; ++p1)
      if (x[p1] >= xn[k + 1]) break;
    p1 = p1 - 1;
    const double khat = fmax(xn[k + 1] - xn[k], xtol);
    double acc = 0.0;
    for (int i = p0; i <= p1; ++i) {
      const double lo = fmax(x[i], xn[k]), hi = fmin(x[i + 1], xn[k + 1]);
      const double mid = x[i + 1] * 0.5 + x[i] * 0.5, half = x[i + 1] * 0.5 - x[i] * 0.5;
      const double sl = (lo - mid) / fmax(half, xtol), sh = (hi - mid) / fmax(half, xtol);
      const double w0 = sh / 1.0 - sl / 1.0, w1 = ppr_p2(sh) / 2.0 - ppr_p2(sl) / 2.0;
      const double w2 = ppr_p3(sh) / 3.0 - ppr_p3(sl) / 3.0;
      double in = w0 * fh[i][0] + w1 * fh[i][1] + w2 * fh[i][2];
      in = in * half;
      acc = acc + in;
    }
    fnew[k] = acc / khat;
    if (fnew[k] > fnew[kmax])
      kmax = k;
    else if (fnew[k] < fnew[kmin])
      kmin = k;
  }
  double so = 0.0, eo = 0.0, sn = 0.0, en = 0.0;
  for (int i = 1; i <= N; ++i) ppr_kbn(&so, &eo, (x[i + 1] - x[i]) * f[i]);
  so = so + eo;
  for (int i = 1; i <= N; ++i) ppr_kbn(&sn, &en, (xn[i + 1] - xn[i]) * fnew[i]);
  sn = sn + en;
  const double err = so - sn;
  if (err > 0.0)
    fnew[kmin] = fnew[kmin] + err / (xn[kmin + 1] - xn[kmin]);
  else if (err < 0.0)
    fnew[kmax] = fnew[kmax] + err / (xn[kmax + 1] - xn[kmax]);
  for (int k = 1; k <= N; ++k) vn[k - 1] = fnew[k];
  free(dx); free(edge); free(fh); free(fnew);
}

void wto_remap_column(int scheme, int N, double* vn, const double* zn, const double* vo, const double* zo) {
  if (scheme == 0)
    remap0(N, vn, zn, vo, zo);
  else if (scheme == 1)
    remap1(N, vn, zn, vo, zo);
  else if (scheme == 2)
    remap2ppm(N, vn, zn, vo, zo);
  else if (scheme == 3)
    remap_ppr(N, vn, zn, vo, zo);
  else if (scheme == 4)
    remap2s(N, vn, zn, vo, zo);
  else if (scheme == 5)
    remap2w(N, vn, zn, vo, zo);
  else
    remap4(N, vn, zn, vo, zo);
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
