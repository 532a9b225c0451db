/* CPU oracle of the vertical diffusion of an ocean column (test
 * infrastructure only; see header).
 *
 * Restates, in the evaluation order of the reference's expressions:
 *   dgtsv (LAPACK tridiagonal solve with partial pivoting, one right-hand
 *          side)                                        dgtsv.f90
 *   turbulent_diffusion: analytic Kt, Kv profiles or the one-equation TKE
 *          closure (init_diffuse, dUdZ_sq, N_sq, l_scales, Richardson,
 *          Prandtl, Kv_tke, Kt_tke, update_Kv_Kt)       vert_diffusion.f90:88-264,
 *                                                       438-573
 *   kinetic_energy (twice the TRiSK kinetic energy of the node's six edges)
 *                                                       io.f90:169-206
 *   backwards_euler_temp (the top row's diagonal takes the super-diagonal
 *          entry of the row below, as the reference writes it)
 *                                                       vert_diffusion.f90:266-355
 *   backwards_euler_velo (Kv of the edge's node)        vert_diffusion.f90:357-424
 *   dz_i, dz_e, free_surface, interp                    utils.f90:135-182, 505-543
 */
#include "vdiff_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define H 2
#define CELL(lz, i, j) ((int64_t)(lz) * S * S + (int64_t)((j) + H) * S + (int64_t)((i) + H))

int wto_dgtsv(int n, double* dl, double* d, double* du, double* b) {
  /* arrays are 0-based here: row i of the library is [i - 1] */
  if (n == 0) return 0;
  for (int i = 1; i <= n - 2; ++i) {
    double* D = d - 1;
    double* DL = dl - 1;
    double* DU = du - 1;
    double* B = b - 1;
    if (fabs(D[i]) >= fabs(DL[i])) {
      if (D[i] != 0.0) {
        const double fact = DL[i] / D[i];
        D[i + 1] = D[i + 1] - fact * DU[i];
        B[i + 1] = B[i + 1] - fact * B[i];
      } else {
        return i;
      }
      DL[i] = 0.0;
    } else {
      const double fact = D[i] / DL[i];
      D[i] = DL[i];
      double temp = D[i + 1];
      D[i + 1] = DU[i] - fact * temp;
      DL[i] = DU[i + 1];
      DU[i + 1] = -fact * DL[i];
      DU[i] = temp;
      temp = B[i];
      B[i] = B[i + 1];
      B[i + 1] = temp - fact * B[i + 1];
    }
  }
  double *D = d - 1, *DL = dl - 1, *DU = du - 1, *B = b - 1;
  if (n > 1) {
    const int i = n - 1;
    if (fabs(D[i]) >= fabs(DL[i])) {
      if (D[i] != 0.0) {
        const double fact = DL[i] / D[i];
        D[i + 1] = D[i + 1] - fact * DU[i];
        B[i + 1] = B[i + 1] - fact * B[i];
      } else {
        return i;
      }
    } else {
      const double fact = D[i] / DL[i];
      D[i] = DL[i];
      double temp = D[i + 1];
      D[i + 1] = DU[i] - fact * temp;
      DU[i] = temp;
      temp = B[i];
      B[i] = B[i + 1];
      B[i + 1] = temp - fact * B[i + 1];
    }
  }
  if (D[n] == 0.0) return n;
  B[n] = B[n] / D[n];
  if (n > 1) B[n - 1] = (B[n - 1] - DU[n - 1] * B[n]) / D[n - 1];
  for (int i = n - 2; i >= 1; --i) B[i] = (B[i] - DU[i] * B[i + 1] - DL[i] * B[i + 2]) / D[i];
  return 0;
}

static double dz_i(const wto_grid* g, const double* mass, int k, int64_t c) {
  return (0.0 + mass[(int64_t)k * g->ncells + c]) / (g->rho0 * 1.0);
}

/* kinetic_energy of layer k at node c (RT, DG, UP of c; RT of W, DG of SW, UP of S) */
static double kinetic_energy(const wto_grid* g, const double* velo, int k, int64_t c) {
  const int64_t NC = g->ncells, S = g->S;
  const double* u = velo + (int64_t)k * 3 * NC;
  const double *len = g->len, *ped = g->pedlen;
  const int64_t W = c - 1, SW = c - S - 1, Sx = c - S;
#define U(e, x) u[(int64_t)(e) * NC + (x)]
#define LEN(e, x) len[(int64_t)(e) * NC + (x)]
#define PED(e, x) ped[(int64_t)(e) * NC + (x)]
  const double pRT = U(0, c) * LEN(0, c), pDG = U(1, c) * LEN(1, c), pUP = U(2, c) * LEN(2, c);
  const double dRT = U(0, c) * PED(0, c), dDG = U(1, c) * PED(1, c), dUP = U(2, c) * PED(2, c);
  const double pUPS = U(2, Sx) * LEN(2, Sx), pDGSW = U(1, SW) * LEN(1, SW), pRTW = U(0, W) * LEN(0, W);
  const double dRTW = U(0, W) * PED(0, W), dDGSW = U(1, SW) * PED(1, SW), dUPS = U(2, Sx) * PED(2, Sx);
#undef U
#undef LEN
#undef PED
  return (pUP * dUP + pDG * dDG + pRT * dRT + pUPS * dUPS + pDGSW * dDGSW + pRTW * dRTW) * g->hex_inv[c] / 4.0;
}

static double prandtl(double Ri) {
  if (Ri < 0.2) return 1.0;
  if (Ri >= 0.2 && Ri <= 2.0) return 5.0 * Ri;
  return 10.0;
}

static double kv_tke(const wto_vdiff* p, double e, double lk) { return fmin(p->Kv_max, fmax(p->C_k * lk * sqrt(e), p->Kv_min)); }

static double kt_tke(const wto_vdiff* p, double Kv, double Nsq, double Ri) {
  double kt = fmin(p->Kt_max, fmax(Kv / prandtl(Ri), p->Kt_min));
  if (p->enhance_diff && Nsq <= p->Nsq_min) kt = p->Kt_enh;
  return kt;
}

/* l_scales (vert_diffusion.f90:502-536): arrays [0..L] */
static void l_scales(const wto_grid* g, const wto_vdiff* p, int L, const double* dz, const double* Nsq, double tau,
                     const double* e, double* l_eps, double* l_k) {
  double* dn = (double*)malloc(sizeof(double) * (L + 1));
  double* up = (double*)malloc(sizeof(double) * (L + 1));
  for (int l = 0; l <= L; ++l) dn[l] = sqrt(2.0 * e[l] / fmax(Nsq[l], p->Neps_sq));
  memcpy(up, dn, sizeof(double) * (L + 1));
  dn[0] = p->l_0;
  for (int l = 1; l <= L; ++l) dn[l] = fmin(dn[l], dn[l - 1] + dz[l]);
  up[L] = fmax(p->l_0, p->kappa_VK * p->C_l / (g->rho0 * g->grav) * tau);
  for (int l = L - 1; l >= 0; --l) up[l] = fmin(up[l], up[l + 1] + dz[l + 1]);
  for (int l = 0; l <= L; ++l) {
    l_eps[l] = fmax(p->l_min, sqrt(up[l] * dn[l]));
    l_k[l] = fmax(p->l_min, fmin(up[l], dn[l]));
  }
  free(dn);
  free(up);
}

static double free_surface(const wto_grid* g, const double* mass, int L, int64_t c) {
  double depth = 0.0;
  for (int k = 0; k < L; ++k) depth = depth + dz_i(g, mass, k, c);
  return depth + g->topo[c];
}

/* turbulent_diffusion of node c (vert_diffusion.f90:88-264) */
static void turbulent(const wto_grid* g, const wto_vdiff* p, int L, double dt, int istep, int64_t c, double tau,
                      double eta, const double* mass, const double* temp, const double* velo, double* Kt, double* Kv,
                      double* tke) {
  const int64_t NC = g->ncells;
#define KT(l) Kt[(int64_t)(l) * NC + c]
#define KV(l) Kv[(int64_t)(l) * NC + c]
  double* w = (double*)calloc(16 * (size_t)(L + 2), sizeof(double));
  double *dz = w, *Umag = w + (L + 2), *dzl = w + 2 * (L + 2), *dU2 = w + 3 * (L + 2), *Nsq = w + 4 * (L + 2);
  double *e = w + 5 * (L + 2), *l_eps = w + 6 * (L + 2), *l_k = w + 7 * (L + 2), *S1 = w + 8 * (L + 2);
  double *S2 = w + 9 * (L + 2), *dg = w + 10 * (L + 2), *dl = w + 11 * (L + 2), *du = w + 12 * (L + 2);
  double* rhs = w + 13 * (L + 2);
  if (!p->tke_closure) {
    double z = g->topo[c];
    const double e20 = -20.0;
    KT(0) = p->Kt_min + p->Kt_max * exp(e20 * (z - eta) / p->mixed_layer);
    KV(0) = p->Kv_min + p->Kv_max * exp(e20 * (z - eta) / p->mixed_layer);
    for (int l = 1; l <= L; ++l) {
      z = z + dz_i(g, mass, l - 1, c);
      KT(l) = p->Kt_min + p->Kt_max * exp(e20 * (z - eta) / p->mixed_layer);
      KV(l) = p->Kv_min + p->Kv_max * exp(e20 * (z - eta) / p->mixed_layer);
    }
    free(w);
    return;
  }
  /* init_diffuse: dz, Umag [1..L], dzl [1..L-1], dUdZ2 and N^2 [0..L] */
  for (int k = 1; k <= L; ++k) {
    dz[k] = dz_i(g, mass, k - 1, c);
    Umag[k] = sqrt(2.0 * kinetic_energy(g, velo, k - 1, c));
  }
  for (int l = 1; l <= L - 1; ++l) dzl[l] = 0.5 * (dz[l] + dz[l + 1]);
  for (int l = 0; l <= L; ++l) {
    if (l == 0 || l == L) {
      dU2[l] = 0.0;
    } else {
      const double q = (Umag[l + 1] - Umag[l]) / dzl[l];
      dU2[l] = q * q;
    }
  }
  for (int l = 0; l <= L; ++l) {
    const int m = l == 0 ? 1 : l == L ? L - 1 : l;
    const double bb = temp[(int64_t)(m - 1) * NC + c] / (0.0 + mass[(int64_t)(m - 1) * NC + c]);
    const double ba = temp[(int64_t)m * NC + c] / (0.0 + mass[(int64_t)m * NC + c]);
    const double dzm = 0.5 * (dz[m] + dz[m + 1]);
    Nsq[l] = g->grav * (ba - bb) / dzm;
  }
  e[0] = p->e_0;
  for (int l = 1; l <= L; ++l) e[l] = fmax(tke[(int64_t)(l - 1) * NC + c], p->e_min);
  l_scales(g, p, L, dz, Nsq, tau, e, l_eps, l_k);
  if (istep == 1)
    for (int l = 0; l <= L; ++l) {
      const double Ri = Nsq[l] / (dU2[l] + p->eps_s);
      KV(l) = kv_tke(p, e[l], l_k[l]);
      KT(l) = kt_tke(p, KV(l), Nsq[l], Ri);
    }
  /* sources and sinks of the TKE equation */
  for (int l = 1; l <= L - 1; ++l) {
    const double turb = e[l] == 0.0 ? 0.0 : p->C_eps * sqrt(e[l]) / l_eps[l];
    S1[l] = KV(l) * dU2[l] - KT(l) * Nsq[l];
    if (p->patankar && S1[l] <= 0.0) {
      S1[l] = KV(l) * dU2[l];
      S2[l] = -turb - KT(l) * Nsq[l] / e[l];
    } else {
      S2[l] = -turb;
    }
  }
  /* backward Euler rows 1..L-1 (coeff reads dzl at the row's own l) */
#define COEFF(l, dzv, kv) (dt * p->C_e * (kv) / (dzl[l] * (dzv)))
  {
    int l = 1;
    du[l] = -COEFF(l, dz[l + 1], 0.5 * (KV(l) + KV(l + 1)));
    dg[l] = 1.0 - du[l] - dt * S2[l];
    rhs[l] = e[l] + dt * S1[l];
    for (l = 2; l <= L - 2; ++l) {
      du[l] = -COEFF(l, dz[l + 1], 0.5 * (KV(l) + KV(l + 1)));
      dl[l - 1] = -COEFF(l, dz[l], 0.5 * (KV(l) + KV(l - 1)));
      dg[l] = 1.0 - (du[l] + dl[l - 1]) - dt * S2[l];
      rhs[l] = e[l] + dt * S1[l];
    }
    l = L - 1;
    dl[l - 1] = -COEFF(l, dz[l], 0.5 * (KV(l) + KV(l - 1)));
    dg[l] = 1.0 - dl[l - 1] - dt * S2[l];
    rhs[l] = e[l] + dt * S1[l];
  }
#undef COEFF
  wto_dgtsv(L - 1, dl + 1, dg + 1, du + 1, rhs + 1);
  e[0] = p->e_0;
  for (int l = 1; l <= L - 1; ++l) {
    e[l] = rhs[l];
    if (!p->patankar) e[l] = fmax(rhs[l], p->e_min);
  }
  e[L] = fmax(p->C_srf * tau / g->rho0, p->e_min_srf);
  /* update_Kv_Kt */
  l_scales(g, p, L, dz, Nsq, tau, e, l_eps, l_k);
  for (int l = 0; l <= L; ++l) {
    const double Ri = Nsq[l] / (dU2[l] + p->eps_s);
    KV(l) = kv_tke(p, e[l], l_k[l]);
    KT(l) = kt_tke(p, KV(l), Nsq[l], Ri);
  }
  for (int l = 1; l <= L; ++l) tke[(int64_t)(l - 1) * NC + c] = fmax(e[l], p->e_min);
#undef KT
#undef KV
  free(w);
}

static double irradiance(const wto_vdiff* p, double depth) {
  return p->Q_sr * (0.58 * exp(-depth / 0.35) + (1.0 - 0.58) * exp(-depth / 23.0));
}

/* backwards_euler_temp of node c (vert_diffusion.f90:266-355) */
static void temp_column(const wto_grid* g, const wto_vdiff* p, int L, double dt, int64_t c, double bot, double top,
                        double eta, const double* Kt, const double* mass, double* temp) {
  const int64_t NC = g->ncells;
  double* w = (double*)calloc(7 * (size_t)(L + 2), sizeof(double));
  double *z = w, *dz = w + (L + 2), *dzl = w + 2 * (L + 2), *dg = w + 3 * (L + 2), *dl = w + 4 * (L + 2);
  double *du = w + 5 * (L + 2), *rhs = w + 6 * (L + 2);
#define KT(l) Kt[(int64_t)(l) * NC + c]
#define B(k) (temp[(int64_t)((k) - 1) * NC + c] / (0.0 + mass[(int64_t)((k) - 1) * NC + c]))
#define SOLAR(k) (p->Q_sr * (irradiance(p, eta - z[k]) - irradiance(p, eta - z[(k) - 1])) / (g->rho0 * p->c_p))
#define COEFF(k, l) (dt * KT((k) + ((l) < 0 ? (l) : 0)) / (dzl[(k) + ((l) < 0 ? (l) : 0)] * dz[k]))
  z[0] = g->topo[c];
  for (int k = 1; k <= L; ++k) {
    dz[k] = dz_i(g, mass, k - 1, c);
    z[k] = z[k - 1] + dz[k];
  }
  for (int l = 1; l <= L - 1; ++l) dzl[l] = 0.5 * (dz[l] + dz[l + 1]);
  int k = 1;
  du[k] = -COEFF(k, 1);
  dg[k] = 1.0 - du[k];
  rhs[k] = B(k) + dt * (-bot + SOLAR(k)) / dz[k];
  for (k = 2; k <= L - 1; ++k) {
    du[k] = -COEFF(k, 1);
    dl[k - 1] = -COEFF(k, -1);
    dg[k] = 1.0 - (du[k] + dl[k - 1]);
    rhs[k] = B(k) + dt * SOLAR(k) / dz[k];
  }
  k = L;
  dl[k - 1] = -COEFF(k, -1);
  dg[k] = 1.0 - du[k - 1];
  rhs[k] = B(k) + dt * (top + p->Q_sr / (g->rho0 * p->c_p)) / dz[k];
  wto_dgtsv(L, dl + 1, dg + 1, du + 1, rhs + 1);
  for (k = 1; k <= L; ++k) {
    const double rho_dz = 0.0 + mass[(int64_t)(k - 1) * NC + c];
    temp[(int64_t)(k - 1) * NC + c] = rho_dz * rhs[k];
  }
#undef KT
#undef B
#undef SOLAR
#undef COEFF
  free(w);
}

/* backwards_euler_velo of node c's three edges (vert_diffusion.f90:357-424) */
static void velo_column(const wto_grid* g, const wto_vdiff* p, int L, double dt, int64_t c, const double* wind,
                        const double* Kv, const double* mass, double* velo) {
  const int64_t NC = g->ncells, S = g->S;
  const int64_t nb[3] = {c + 1, c + S + 1, c + S};
  double* w = (double*)calloc(6 * (size_t)(L + 2), sizeof(double));
  double *dz = w, *dzl = w + (L + 2), *dg = w + 2 * (L + 2), *dl = w + 3 * (L + 2), *du = w + 4 * (L + 2);
  double* r = w + 5 * (L + 2);
#define KV(l) Kv[(int64_t)(l) * NC + c]
#define COEFF(k, l) (dt * KV((k) + ((l) < 0 ? (l) : 0)) / (dzl[(k) + ((l) < 0 ? (l) : 0)] * dz[k]))
  for (int e = 0; e < 3; ++e) {
    for (int k = 1; k <= L; ++k) dz[k] = 0.5 * (dz_i(g, mass, k - 1, c) + dz_i(g, mass, k - 1, nb[e]));
    for (int l = 1; l <= L - 1; ++l) dzl[l] = 0.5 * (dz[l] + dz[l + 1]);
    int k = 1;
    du[k] = -COEFF(k, 1);
    dg[k] = 1.0 - du[k] + dt * p->bottom_friction / dz[k];
    r[k] = velo[((int64_t)(k - 1) * 3 + e) * NC + c];
    for (k = 2; k <= L - 1; ++k) {
      du[k] = -COEFF(k, 1);
      dl[k - 1] = -COEFF(k, -1);
      dg[k] = 1.0 - (du[k] + dl[k - 1]);
      r[k] = velo[((int64_t)(k - 1) * 3 + e) * NC + c];
    }
    k = L;
    dl[k - 1] = -COEFF(k, -1);
    dg[k] = 1.0 - dl[k - 1];
    r[k] = velo[((int64_t)(k - 1) * 3 + e) * NC + c] + dt * wind[(int64_t)e * NC + c] / dz[k];
    wto_dgtsv(L, dl + 1, dg + 1, du + 1, r + 1);
    for (k = 1; k <= L; ++k) velo[((int64_t)(k - 1) * 3 + e) * NC + c] = r[k];
  }
#undef KV
#undef COEFF
  free(w);
}

void wto_vertical_diffusion(const wto_grid* g, const wto_vdiff* p, int L, double dt, int istep, const double* bot,
                            const double* top, const double* tau, const double* wind, const double* eta, double* Kt,
                            double* Kv, double* tke, double* mass, double* temp, double* velo) {
  const int n = g->n, S = g->S;
  const int64_t nown = (int64_t)10 * n * n;
  /* columns: the owned nodes, then the two pole masters */
  int64_t* col = (int64_t*)malloc(sizeof(int64_t) * (nown + 2));
  int64_t q = 0;
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) col[q++] = CELL(lz, i, j);
  col[q++] = g->south_pole;
  col[q++] = g->north_pole;
  for (int64_t x = 0; x < q; ++x) {
    const int64_t c = col[x];
    const double ec = eta ? eta[c] : free_surface(g, mass, L, c);
    turbulent(g, p, L, dt, istep, c, tau[c], ec, mass, temp, velo, Kt, Kv, tke);
  }
  for (int64_t x = 0; x < q; ++x) {
    const int64_t c = col[x];
    const double ec = eta ? eta[c] : free_surface(g, mass, L, c);
    temp_column(g, p, L, dt, c, bot[c], top[c], ec, Kt, mass, temp);
  }
  for (int64_t x = 0; x < nown; ++x) velo_column(g, p, L, dt, col[x], wind, Kv, mass, velo);
  free(col);
  wto_ghost_fill(g, L, mass, temp, velo);
}
