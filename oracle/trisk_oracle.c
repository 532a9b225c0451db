/* CPU oracle of the TRiSK hot path (test infrastructure only; see header).
 *
 * Restates, per uniform level and in each lozenge's own (i, j) frame:
 *   cal_surf_press / column_mass / set_surf_geopot   ops.f90:1023-1092
 *   integrate_pressure_up (incompressible branch)    ops.f90:1094-1131
 *   step1 itype 0: circulations, PV, qe, KE, Bernoulli, Exner, fluxes
 *                                                    ops.f90:276-366, 472-522
 *   post_step1 (pentagon corners)                    ops.f90:526-737
 *   scalar_trend / div                               ops.f90:739-755, 1325-1343
 *   du_source / Qperp / get_weights                  ops.f90:757-785, 834-901, 1006-1021
 *   du_grad / gradi_e                                ops.f90:787-832, 1280-1301
 *   RK4, RK_sub_step                                 time_integr.f90:78-103, 206-225
 *   Laplacian diffusion / hyperdiffusion (wto_trend_diff):
 *     cal_divu_ml, cal_Laplacian_scalars, cal_Laplacian_divu
 *                                                    multi_level.f90:26-30, 190-309, 579-617
 *     cal_Laplacian_vector_rot, cal_Laplacian_rotu, post_vort
 *                                                    multi_level.f90:99, 240-270, ops.f90:1159-1278
 *     step1 itypes 1, 5, 8 and the vorticity of itype 0  ops.f90:196-201, 236-239, 256-275, 334-338
 *     the physics hooks of the reference's test cases (scalar flux
 *     (-1)^Ls visc grad(q or Laplacian q) pedlen, velocity source
 *     (-1)^(Lr-1) (visc_divu grad divu - visc_rotu curl rotu))
 * with the evaluation order of the reference's expressions (compile with
 * -ffp-contract=off).  Halo-edge fluxes use the owning edge's dual length,
 * as the reference's ghost fill of horiz_flux does (multi_level.f90:111).
 */
#include "trisk_oracle.h"

#include <stdlib.h>
#include <string.h>

#define H 2
#define CELL(lz, i, j) ((int64_t)(lz) * S * S + (int64_t)((j) + H) * S + (int64_t)((i) + H))

/* The cell loops of one pass write distinct cells and read only the inputs of
 * that pass, so they run over the host's cores (OpenMP) with results that do
 * not depend on the thread count. */
#ifdef _OPENMP
#include <omp.h>
#define WTO_PARALLEL_LZ_J _Pragma("omp parallel for collapse(2) schedule(static)")
#define WTO_PARALLEL_K_C _Pragma("omp parallel for collapse(2) schedule(static)")
int wto_num_threads(void) { return omp_get_max_threads(); }
#else
#define WTO_PARALLEL_LZ_J
#define WTO_PARALLEL_K_C
int wto_num_threads(void) { return 1; }
#endif

static inline double interp(double a, double b) { return 0.5 * (a + b); }

void wto_ghost_fill(const wto_grid* g, int L, double* mass, double* temp, double* velo) {
  const int64_t NC = g->ncells;
  /* a halo slot's source is an owned slot, so the copies are independent */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int k = 0; k < L; ++k) {
    double* m = mass + (int64_t)k * NC;
    double* t = temp + (int64_t)k * NC;
    for (int64_t c = 0; c < NC; ++c) {
      int64_t s = g->node_src[c];
      if (s >= 0 && s != c) {
        m[c] = m[s];
        t[c] = t[s];
      }
    }
    double* u = velo + (int64_t)k * 3 * NC;
    for (int64_t c = 0; c < NC; ++c)
      for (int e = 0; e < 3; ++e) {
        int64_t slot = 3 * c + e, s = g->edge_src[slot];
        if (s >= 0 && s != slot) u[e * NC + c] = (double)g->edge_sign[slot] * u[(s % 3) * NC + s / 3];
      }
  }
}

/* get_weights (ops.f90:1006-1021) */
static void weights(const wto_grid* g, int64_t c, int offs, double w[5]) {
  const int64_t NC = g->ncells;
  double acc[5];
  acc[0] = g->area[(int64_t)offs * NC + c];
  for (int i = 2; i <= 5; ++i) acc[i - 1] = acc[i - 2] + g->area[(int64_t)((i + offs - 1) % 6) * NC + c];
  for (int i = 0; i < 5; ++i) acc[i] = 0.5 - acc[i] * g->hex_inv[c];
  w[0] = acc[0];
  w[1] = -acc[1];
  w[2] = acc[2];
  w[3] = -acc[3];
  w[4] = acc[4];
}

typedef struct {
  double *pvL, *pvU, *qe, *hm, *ht, *ke, *bern, *exner;
  double *press, *press_lower, *geopot, *geopot_lower, *ped_flux;
  /* diffusion: owner's edge length of halo edges, divu and its Laplacian,
   * the scalars' Laplacians, the vorticity (or rot rot of it) on LORT / UPLT
   * of every cell, rot of the vorticity on edges */
  double *len_flux, *divu, *lapdivu, *lapm, *lapt, *vortL, *vortU, *rotu;
} work_t;

/* node field of one layer: owners' values into the halo */
static void node_ghost(const wto_grid* g, double* f) {
  for (int64_t c = 0; c < g->ncells; ++c) {
    int64_t s = g->node_src[c];
    if (s >= 0 && s != c) f[c] = f[s];
  }
}

static void edge_ghost(const wto_grid* g, double* u) {
  const int64_t NC = g->ncells;
  for (int64_t c = 0; c < NC; ++c)
    for (int e = 0; e < 3; ++e) {
      int64_t slot = 3 * c + e, s = g->edge_src[slot];
      if (s >= 0 && s != slot) u[e * NC + c] = (double)g->edge_sign[slot] * u[(s % 3) * NC + s / 3];
    }
}

static int is_master_cell(const wto_grid* g, int64_t c);

/* div (ops.f90:1325-1343) of the edge field F on the masters, then the halo:
 * cal_div followed by the ghost fill */
static void div_masters(const wto_grid* g, const double* F, double* out) {
  const int S = g->S;
  const int64_t NC = g->ncells;
  for (int64_t c = 0; c < NC; ++c) {
    if (!is_master_cell(g, c)) continue;
    const int64_t cW = c - 1, cS = c - S, cSW = c - S - 1;
    out[c] = (F[c] - F[cW] + F[NC + cSW] - F[NC + c] + F[2 * NC + c] - F[2 * NC + cS]) * g->hex_inv[c];
  }
  node_ghost(g, out);
}

/* div of the edge field F as cal_divu_ml leaves it (multi_level.f90:608-614):
 * cal_div over i, j in [0, n] of every lozenge with no ghost fill after it,
 * so the halo row and column i = n or j = n hold each lozenge's own
 * divergence of its ghost fluxes.  Away from the corner (n, n) that equals the
 * owner's value; at that pentagon corner it does not.  The rest of the halo
 * takes the owners' values (it is never read). */
static void div_local(const wto_grid* g, const double* F, double* out) {
  const int n = g->n, S = g->S;
  const int64_t NC = g->ncells;
  div_masters(g, F, out);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i) {
        if (i < n && j < n) continue;
        int64_t c = CELL(lz, i, j);
        const int64_t cW = c - 1, cS = c - S, cSW = c - S - 1;
        out[c] = (F[c] - F[cW] + F[NC + cSW] - F[NC + c] + F[2 * NC + c] - F[2 * NC + cS]) * g->hex_inv[c];
      }
}

/* Laplacian of a node field (step1 itype 1, then cal_div): gradient flux
 * (q(nb) - q) / len * pedlen with the owner's lengths on halo edges */
static void laplacian_nodes(const wto_grid* g, const double* lenf, const double* pedf, const double* q, double* F,
                            double* out, int plant) {
  const int n = g->n, S = g->S;
  const int64_t NC = g->ncells;
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -1; j <= n; ++j)
      for (int i = -1; i <= n; ++i) {
        int64_t c = CELL(lz, i, j);
        F[c] = (q[CELL(lz, i + 1, j)] - q[c]) / lenf[c] * pedf[c];
        F[NC + c] = (q[c] - q[CELL(lz, i + 1, j + 1)]) / lenf[NC + c] * pedf[NC + c];
        F[2 * NC + c] = (q[CELL(lz, i, j + 1)] - q[c]) / lenf[2 * NC + c] * pedf[2 * NC + c];
      }
  /* update_bdry of horiz_flux: halo edges take their owners' fluxes, which
   * differ from the local ones where q holds a lozenge's own corner value */
  if (!plant) edge_ghost(g, F);
  div_masters(g, F, out);
}

/* circulation / area of every LORT and UPLT triangle of cells [-1, n]^2 of
 * edge field u, and the pentagon corners' merged triangles (post_step1,
 * post_vort) */
static void vorticity(const wto_grid* g, const double* u, double* vL, double* vU) {
  const int n = g->n, S = g->S;
  const int64_t NC = g->ncells;
  const double *len = g->len, *tri = g->triarea;
#define UV(e, c) (u[(int64_t)(e) * NC + (c)] * len[(int64_t)(e) * NC + (c)])
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -1; j <= n; ++j)
      for (int i = -1; i <= n; ++i) {
        int64_t c = CELL(lz, i, j), cE = CELL(lz, i + 1, j), cN = CELL(lz, i, j + 1);
        vL[c] = (UV(0, c) + UV(2, cE) + UV(1, c)) / tri[c];
        vU[c] = -(UV(1, c) + UV(2, c) + UV(0, cN)) / tri[NC + c];
      }
  for (int lz = 0; lz < 10; ++lz) {
    { /* IJMINUS */
      int64_t idS = CELL(lz, 0, -1), idW = CELL(lz, -1, 0), idSW = CELL(lz, -1, -1);
      vL[idSW] = (UV(2, idS) - UV(0, idW) + UV(0, idSW)) / tri[idSW];
      vU[idSW] = vL[idSW];
    }
    { /* IPLUSJMINUS */
      int64_t id = CELL(lz, n, 0), idS = CELL(lz, n, -1), idSW = CELL(lz, n - 1, -1);
      vL[idSW] = (-UV(0, id) + UV(0, idSW) + UV(1, idSW)) / tri[idSW];
      vU[idS] = vL[idSW];
    }
    { /* IMINUSJPLUS */
      int64_t id = CELL(lz, 0, n), idW = CELL(lz, -1, n), idSW = CELL(lz, -1, n - 1);
      vU[idSW] = (UV(2, id) - UV(1, idSW) - UV(2, idSW)) / tri[NC + idSW];
      vL[idW] = vU[idSW];
    }
    { /* IJPLUS */
      int64_t id = CELL(lz, n, n), idN = CELL(lz, n, n + 1);
      vL[id] = (UV(0, id) - UV(0, idN) - UV(2, id)) / tri[id];
      vU[id] = vL[id];
    }
  }
#undef UV
}

static void trend_impl(const wto_grid* g, int L, const double* mass, const double* temp, const double* velo,
                       double* dmass, double* dtemp, double* dvelo, wto_trend_keep* keep, const wto_diff* dif);

void wto_trend(const wto_grid* g, int L, const double* mass, const double* temp, const double* velo,
               double* dmass, double* dtemp, double* dvelo) {
  trend_impl(g, L, mass, temp, velo, dmass, dtemp, dvelo, NULL, NULL);
}

void wto_trend_parts(const wto_grid* g, int L, const double* mass, const double* temp, const double* velo,
                     double* dmass, double* dtemp, double* dvelo, wto_trend_keep* keep) {
  trend_impl(g, L, mass, temp, velo, dmass, dtemp, dvelo, keep, NULL);
}

void wto_trend_diff(const wto_grid* g, const wto_diff* dif, int L, const double* mass, const double* temp,
                    const double* velo, double* dmass, double* dtemp, double* dvelo) {
  trend_impl(g, L, mass, temp, velo, dmass, dtemp, dvelo, NULL, dif);
}

static void trend_impl(const wto_grid* g, int L, const double* mass, const double* temp, const double* velo,
                       double* dmass, double* dtemp, double* dvelo, wto_trend_keep* keep, const wto_diff* dif) {
  const int n = g->n, S = g->S;
  const int64_t NC = g->ncells;
  const double grav = g->grav, rho0 = g->rho0;
  const double *len = g->len, *ped = g->pedlen, *A = g->area, *hinv = g->hex_inv;
  const double *tri = g->triarea, *cor = g->coriolis;
  work_t w;
  double** ptrs[] = {&w.pvL, &w.pvU, &w.ke, &w.bern, &w.exner, &w.press, &w.press_lower, &w.geopot, &w.geopot_lower};
  for (unsigned q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) *ptrs[q] = (double*)calloc(NC, sizeof(double));
  w.qe = (double*)calloc(3 * NC, sizeof(double));
  w.hm = (double*)calloc(3 * NC, sizeof(double));
  w.ht = (double*)calloc(3 * NC, sizeof(double));
  w.ped_flux = (double*)calloc(3 * NC, sizeof(double));
  const int dsc = dif ? dif->lsclr : 0, ddv = dif ? dif->ldivu : 0, drt = dif ? dif->lrotu : 0;
  double* gflux = NULL; /* gradient fluxes of the Laplacians */
  w.len_flux = w.divu = w.lapdivu = w.lapm = w.lapt = w.vortL = w.vortU = w.rotu = NULL;
  if (dif) {
    w.len_flux = (double*)calloc(3 * NC, sizeof(double));
    w.divu = (double*)calloc(NC, sizeof(double));
    w.lapdivu = (double*)calloc(NC, sizeof(double));
    w.lapm = (double*)calloc(NC, sizeof(double));
    w.lapt = (double*)calloc(NC, sizeof(double));
    w.vortL = (double*)calloc(NC, sizeof(double));
    w.vortU = (double*)calloc(NC, sizeof(double));
    w.rotu = (double*)calloc(3 * NC, sizeof(double));
    gflux = (double*)calloc(3 * NC, sizeof(double));
  }
  memset(dmass, 0, sizeof(double) * NC * L);
  memset(dtemp, 0, sizeof(double) * NC * L);
  memset(dvelo, 0, sizeof(double) * 3 * NC * L);

  /* dual lengths used for fluxes: a halo edge carries its owner's flux */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t c = 0; c < NC; ++c)
    for (int e = 0; e < 3; ++e) {
      int64_t slot = 3 * c + e, s = g->edge_src[slot];
      w.ped_flux[e * NC + c] = (s >= 0 && s != slot) ? ped[(s % 3) * NC + s / 3] : ped[e * NC + c];
      if (w.len_flux) w.len_flux[e * NC + c] = (s >= 0 && s != slot) ? len[(s % 3) * NC + s / 3] : len[e * NC + c];
    }

  /* surface pressure and geopotential (cal_surf_press) on [0, n]^2 */
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -1; j <= n + 1; ++j)
      for (int i = -1; i <= n + 1; ++i) w.geopot[CELL(lz, i, j)] = grav * g->topo[CELL(lz, i, j)];
  WTO_PARALLEL_LZ_J
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i) {
        int64_t c = CELL(lz, i, j);
        double sp = 0.0;
        for (int k = 0; k < L; ++k) sp = sp + (mass[(int64_t)k * NC + c] - temp[(int64_t)k * NC + c]);
        w.press_lower[c] = grav * sp + g->p_top;
      }

  for (int k = 0; k < L; ++k) {
    const double* m = mass + (int64_t)k * NC;
    const double* t = temp + (int64_t)k * NC;
    const double* u = velo + (int64_t)k * 3 * NC;
    double* dm = dmass + (int64_t)k * NC;
    double* dt = dtemp + (int64_t)k * NC;
    double* du = dvelo + (int64_t)k * 3 * NC;
#define U(e, c) (u[(int64_t)(e) * NC + (c)])
#define LEN(e, c) (len[(int64_t)(e) * NC + (c)])
#define PED(e, c) (ped[(int64_t)(e) * NC + (c)])
#define AR(p, c) (A[(int64_t)((p) - 1) * NC + (c)])
#define UP_(e, c) (U(e, c) * LEN(e, c))
#define UD_(e, c) (U(e, c) * PED(e, c))
    /* integrate_pressure_up on [0, n]^2 */
    WTO_PARALLEL_LZ_J
    for (int lz = 0; lz < 10; ++lz)
      for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
          int64_t c = CELL(lz, i, j);
          double rho_dz = m[c], rho_dz_theta = t[c];
          w.geopot_lower[c] = w.geopot[c];
          double p_upper = w.press_lower[c] - grav * (rho_dz - rho_dz_theta);
          w.press[c] = interp(w.press_lower[c], p_upper);
          w.geopot[c] = w.geopot_lower[c] + grav * rho_dz / (rho0 * 1.0);
          w.press_lower[c] = p_upper;
        }
    /* diffusion (the order of trend_ml, multi_level.f90:26-30): divu of the
     * velocity and its Laplacian, the scalars' Laplacians, the vorticity and
     * (order 2) rot rot of it (multi_level.f90:99) */
    if (ddv) {
      for (int64_t c = 0; c < NC; ++c)
        for (int e = 0; e < 3; ++e) gflux[e * NC + c] = U(e, c) * w.ped_flux[e * NC + c];
      if (dif->plant_owner_halo)
        div_masters(g, gflux, w.divu);
      else
        div_local(g, gflux, w.divu);
      if (ddv == 2) laplacian_nodes(g, w.len_flux, w.ped_flux, w.divu, gflux, w.lapdivu, dif->plant_owner_halo);
    }
    if (dsc == 2) {
      laplacian_nodes(g, w.len_flux, w.ped_flux, m, gflux, w.lapm, dif->plant_owner_halo);
      laplacian_nodes(g, w.len_flux, w.ped_flux, t, gflux, w.lapt, dif->plant_owner_halo);
    }
    if (drt) {
      vorticity(g, u, w.vortL, w.vortU);
      if (drt == 2) {
        memset(w.rotu, 0, sizeof(double) * 3 * NC);
        for (int lz = 0; lz < 10; ++lz)
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
              int64_t c = CELL(lz, i, j), cS = CELL(lz, i, j - 1), cW = CELL(lz, i - 1, j);
              w.rotu[c] = -(w.vortL[c] - w.vortU[cS]) / PED(0, c);
              w.rotu[NC + c] = PED(1, c) != 0.0 ? -(w.vortL[c] - w.vortU[c]) / PED(1, c) : 0.0;
              w.rotu[2 * NC + c] = -(w.vortL[cW] - w.vortU[c]) / PED(2, c);
            }
        edge_ghost(g, w.rotu);
        vorticity(g, w.rotu, w.vortL, w.vortU);
      }
    }
    /* potential vorticity on triangles LORT(i,j), UPLT(i,j), (i,j) in [-1, n]^2 */
    WTO_PARALLEL_LZ_J
    for (int lz = 0; lz < 10; ++lz)
      for (int j = -1; j <= n; ++j)
        for (int i = -1; i <= n; ++i) {
          int64_t c = CELL(lz, i, j), cE = CELL(lz, i + 1, j), cN = CELL(lz, i, j + 1), cNE = CELL(lz, i + 1, j + 1);
          double circL = UP_(0, c) + UP_(2, cE) + UP_(1, c);
          double circU = -(UP_(1, c) + UP_(2, c) + UP_(0, cN));
          double massL = m[c] * AR(1, c) + m[cE] * AR(3, cE) + m[cNE] * AR(5, cNE);
          double massU = m[c] * AR(2, c) + m[cNE] * AR(4, cNE) + m[cN] * AR(6, cN);
          w.pvL[c] = (cor[c] + circL) / massL;
          w.pvU[c] = (cor[NC + c] + circU) / massU;
        }
    /* qe and fluxes on every edge a stencil reads */
    WTO_PARALLEL_LZ_J
    for (int lz = 0; lz < 10; ++lz)
      for (int j = -1; j <= n; ++j)
        for (int i = -1; i <= n; ++i) {
          int64_t c = CELL(lz, i, j);
          int64_t nb[3] = {CELL(lz, i + 1, j), CELL(lz, i + 1, j + 1), CELL(lz, i, j + 1)};
          if (j >= 0) w.qe[c] = interp(w.pvU[CELL(lz, i, j - 1)], w.pvL[c]);
          w.qe[NC + c] = interp(w.pvU[c], w.pvL[c]);
          if (i >= 0) w.qe[2 * NC + c] = interp(w.pvU[c], w.pvL[CELL(lz, i - 1, j)]);
          for (int e = 0; e < 3; ++e) {
            double udual = U(e, c) * w.ped_flux[e * NC + c];
            double pm = 0.0, pt = 0.0;
            if (dsc) { /* physics_scalar_flux: (-1)^Ls visc grad(q or Laplacian q) pedlen */
              const double* qm = dsc == 1 ? m : w.lapm;
              const double* qt = dsc == 1 ? t : w.lapt;
              const double sg = (dsc % 2) ? -1.0 : 1.0;
              const double dl = w.len_flux[e * NC + c], pl = w.ped_flux[e * NC + c];
              const double gm = e == 1 ? (qm[c] - qm[nb[e]]) / dl : (qm[nb[e]] - qm[c]) / dl;
              const double gt = e == 1 ? (qt[c] - qt[nb[e]]) / dl : (qt[nb[e]] - qt[c]) / dl;
              pm = sg * dif->visc_m * gm * pl;
              pt = sg * dif->visc_t * gt * pl;
            }
            w.hm[e * NC + c] = udual * interp(m[c], m[nb[e]]) + pm;
            w.ht[e * NC + c] = udual * interp(t[c], t[nb[e]]) + pt;
          }
        }
    /* pentagon corners (post_step1) */
    for (int lz = 0; lz < 10; ++lz) {
      { /* IJMINUS: lower-left corner (0,0) */
        int64_t id = CELL(lz, 0, 0), idN = CELL(lz, 0, 1), idE = CELL(lz, 1, 0), idS = CELL(lz, 0, -1);
        int64_t idW = CELL(lz, -1, 0), idSW = CELL(lz, -1, -1);
        double cSW_L = UP_(2, idS) - UP_(0, idW) + UP_(0, idSW);
        double cW_L = ((UP_(0, idW) + UP_(2, id) + UP_(1, idW)) / tri[idW]) * tri[idW];
        double cS_U = (-(UP_(0, id) + UP_(1, idS) + UP_(2, idS)) / tri[NC + idS]) * tri[NC + idS];
        double m1 = m[idW] * AR(6, idW) + m[id] * (AR(4, id) + AR(5, id)) + m[idS] * AR(3, idS);
        double m2 = m[idW] * AR(1, idW) + m[id] * AR(3, id) + m[idN] * AR(5, idN);
        double m3 = m[idS] * AR(2, idS) + m[idE] * AR(4, idE) + m[id] * AR(6, id);
        double pvWL = (cor[idW] + cW_L) / m1;
        double pvSWL = (cor[idSW] + cSW_L) / m2;
        double pvSU = (cor[NC + idS] + cS_U) / m3;
        w.qe[idW] = interp(pvWL, pvSWL);
        w.qe[2 * NC + idS] = interp(pvSU, pvSWL);
      }
      { /* IPLUSJMINUS: lower-right corner (n,0) */
        int64_t id = CELL(lz, n, 0), idE = CELL(lz, n + 1, 0), idS = CELL(lz, n, -1), idW = CELL(lz, n - 1, 0);
        int64_t idNE = CELL(lz, n + 1, 1), idSW = CELL(lz, n - 1, -1);
        double cSW_L = -UP_(0, id) + UP_(0, idSW) + UP_(1, idSW);
        double cL = ((UP_(0, id) + UP_(2, idE) + UP_(1, id)) / tri[id]) * tri[id];
        double cSW_U = (-(UP_(0, idW) + UP_(1, idSW) + UP_(2, idSW)) / tri[NC + idSW]) * tri[NC + idSW];
        double m1 = m[idSW] * AR(1, idSW) + m[idS] * AR(3, idS) + m[id] * (AR(5, id) + AR(6, id));
        double m2 = m[id] * AR(1, id) + m[idE] * AR(3, idE) + m[idNE] * AR(5, idNE);
        double m3 = m[idSW] * AR(2, idSW) + m[id] * AR(4, id) + m[idW] * AR(6, idW);
        double pvSWL = (cor[idSW] + cSW_L) / m1;
        double pvL = (cor[id] + cL) / m2;
        double pvSWU = (cor[NC + idSW] + cSW_U) / m3;
        w.qe[id] = interp(pvSWL, pvL);
        w.qe[NC + idSW] = interp(pvSWL, pvSWU);
      }
      { /* IMINUSJPLUS: upper-left corner (0,n) */
        int64_t id = CELL(lz, 0, n), idN = CELL(lz, 0, n + 1), idS = CELL(lz, 0, n - 1), idW = CELL(lz, -1, n);
        int64_t idNE = CELL(lz, 1, n + 1), idSW = CELL(lz, -1, n - 1);
        double cSW_U = UP_(2, id) - UP_(1, idSW) - UP_(2, idSW);
        double cU = (-(UP_(1, id) + UP_(2, id) + UP_(0, idN)) / tri[NC + id]) * tri[NC + id];
        double cSW_L = ((UP_(0, idSW) + UP_(2, idS) + UP_(1, idSW)) / tri[idSW]) * tri[idSW];
        double m1 = m[idSW] * AR(2, idSW) + m[id] * (AR(3, id) + AR(4, id)) + m[idW] * AR(6, idW);
        double m2 = m[id] * AR(2, id) + m[idNE] * AR(4, idNE) + m[idN] * AR(6, idN);
        double m3 = m[idSW] * AR(1, idSW) + m[idS] * AR(3, idS) + m[id] * AR(5, id);
        double pvSWU = (cor[NC + idSW] + cSW_U) / m1;
        double pvU = (cor[NC + id] + cU) / m2;
        double pvSWL = (cor[idSW] + cSW_L) / m3;
        w.qe[2 * NC + id] = interp(pvSWU, pvU);
        w.qe[NC + idSW] = interp(pvSWL, pvSWU);
      }
      { /* IJPLUS: upper-right corner (n,n) */
        int64_t id = CELL(lz, n, n), idN = CELL(lz, n, n + 1), idE = CELL(lz, n + 1, n), idS = CELL(lz, n, n - 1);
        int64_t idW = CELL(lz, n - 1, n);
        double cL = UP_(0, id) - UP_(0, idN) - UP_(2, id);
        double cW_L = ((UP_(0, idW) + UP_(2, id) + UP_(1, idW)) / tri[idW]) * tri[idW];
        double cS_U = (-(UP_(1, idS) + UP_(2, idS) + UP_(0, id)) / tri[NC + idS]) * tri[NC + idS];
        double m1 = m[idE] * AR(3, idE) + m[id] * (AR(1, id) + AR(2, id)) + m[idN] * AR(6, idN);
        double m2 = m[idW] * AR(1, idW) + m[id] * AR(3, id) + m[idN] * AR(5, idN);
        double m3 = m[idS] * AR(2, idS) + m[idE] * AR(4, idE) + m[id] * AR(6, id);
        double pvL = (cor[id] + cL) / m1;
        double pvWL = (cor[idW] + cW_L) / m2;
        double pvSU = (cor[NC + idS] + cS_U) / m3;
        w.qe[id] = interp(pvL, pvSU);
        w.qe[2 * NC + id] = interp(pvL, pvWL);
      }
    }
    /* kinetic energy, Bernoulli, Exner on [0, n]^2 */
    WTO_PARALLEL_LZ_J
    for (int lz = 0; lz < 10; ++lz)
      for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
          int64_t c = CELL(lz, i, j), cW = CELL(lz, i - 1, j), cS = CELL(lz, i, j - 1), cSW = CELL(lz, i - 1, j - 1);
          double ke = (UP_(2, c) * UD_(2, c) + UP_(1, c) * UD_(1, c) + UP_(0, c) * UD_(0, c) +
                       UP_(2, cS) * UD_(2, cS) + UP_(1, cSW) * UD_(1, cSW) + UP_(0, cW) * UD_(0, cW)) *
                      hinv[c] / 4.0;
          double phi_k = interp(w.geopot[c], w.geopot_lower[c]);
          w.ke[c] = ke;
          w.bern[c] = ke + phi_k + w.press[c] / (rho0 * 1.0);
          w.exner[c] = -phi_k;
        }
    if (keep) {
      memcpy(keep->hflux_m + (int64_t)k * 3 * NC, w.hm, sizeof(double) * 3 * NC);
      memcpy(keep->hflux_t + (int64_t)k * 3 * NC, w.ht, sizeof(double) * 3 * NC);
      memcpy(keep->bern + (int64_t)k * NC, w.bern, sizeof(double) * NC);
      memcpy(keep->exner + (int64_t)k * NC, w.exner, sizeof(double) * NC);
    }
    /* scalar trends: -div(flux) on owned cells and the pole masters */
    WTO_PARALLEL_LZ_J
    for (int lz = 0; lz < 10; ++lz)
      for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
          int64_t c = CELL(lz, i, j);
          int owned = (i < n && j < n) || c == g->south_pole || c == g->north_pole;
          if (!owned) continue;
          int64_t cW = CELL(lz, i - 1, j), cS = CELL(lz, i, j - 1), cSW = CELL(lz, i - 1, j - 1);
          double divm = (w.hm[c] - w.hm[cW] + w.hm[NC + cSW] - w.hm[NC + c] + w.hm[2 * NC + c] - w.hm[2 * NC + cS]) * hinv[c];
          double divt = (w.ht[c] - w.ht[cW] + w.ht[NC + cSW] - w.ht[NC + c] + w.ht[2 * NC + c] - w.ht[2 * NC + cS]) * hinv[c];
          dm[c] = -divm;
          dt[c] = -divt;
        }
    /* velocity trends on owned edges: -Qperp, then gradients (du_source, du_grad) */
#define HM(e, c) (w.hm[(int64_t)(e) * NC + (c)])
#define QE(e, c) (w.qe[(int64_t)(e) * NC + (c)])
    WTO_PARALLEL_LZ_J
    for (int lz = 0; lz < 10; ++lz)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          int64_t id = CELL(lz, i, j);
          int64_t idNW = CELL(lz, i - 1, j + 1), idN = CELL(lz, i, j + 1), idNE = CELL(lz, i + 1, j + 1);
          int64_t idW = CELL(lz, i - 1, j), idE = CELL(lz, i + 1, j), idSW = CELL(lz, i - 1, j - 1);
          int64_t idS = CELL(lz, i, j - 1), idSE = CELL(lz, i + 1, j - 1);
          double w1[5], w2[5], q[3];
          weights(g, id, 0, w1);
          weights(g, idE, 3, w2);
          double qr = QE(0, id);
          q[0] = HM(1, id) * interp(QE(1, id), qr) * w1[0] + HM(2, id) * interp(QE(2, id), qr) * w1[1] +
                 HM(0, idW) * interp(QE(0, idW), qr) * w1[2] + HM(1, idSW) * interp(QE(1, idSW), qr) * w1[3] +
                 HM(2, idS) * interp(QE(2, idS), qr) * w1[4] + HM(1, idS) * interp(QE(1, idS), qr) * w2[0] +
                 HM(2, idSE) * interp(QE(2, idSE), qr) * w2[1] + HM(0, idE) * interp(QE(0, idE), qr) * w2[2] +
                 HM(1, idE) * interp(QE(1, idE), qr) * w2[3] + HM(2, idE) * interp(QE(2, idE), qr) * w2[4];
          weights(g, id, 1, w1);
          weights(g, idNE, 4, w2);
          double qd = QE(1, id);
          q[1] = HM(2, id) * interp(QE(2, id), qd) * w1[0] + HM(0, idW) * interp(QE(0, idW), qd) * w1[1] +
                 HM(1, idSW) * interp(QE(1, idSW), qd) * w1[2] + HM(2, idS) * interp(QE(2, idS), qd) * w1[3] +
                 HM(0, id) * interp(QE(0, id), qd) * w1[4] + HM(2, idE) * interp(QE(2, idE), qd) * w2[0] +
                 HM(0, idNE) * interp(QE(0, idNE), qd) * w2[1] + HM(1, idNE) * interp(QE(1, idNE), qd) * w2[2] +
                 HM(2, idNE) * interp(QE(2, idNE), qd) * w2[3] + HM(0, idN) * interp(QE(0, idN), qd) * w2[4];
          weights(g, id, 2, w1);
          weights(g, idN, 5, w2);
          double qu = QE(2, id);
          q[2] = HM(0, idW) * interp(QE(0, idW), qu) * w1[0] + HM(1, idSW) * interp(QE(1, idSW), qu) * w1[1] +
                 HM(2, idS) * interp(QE(2, idS), qu) * w1[2] + HM(0, id) * interp(QE(0, id), qu) * w1[3] +
                 HM(1, id) * interp(QE(1, id), qu) * w1[4] + HM(0, idN) * interp(QE(0, idN), qu) * w2[0] +
                 HM(1, idN) * interp(QE(1, idN), qu) * w2[1] + HM(2, idN) * interp(QE(2, idN), qu) * w2[2] +
                 HM(0, idNW) * interp(QE(0, idNW), qu) * w2[3] + HM(1, idW) * interp(QE(1, idW), qu) * w2[4];
          /* du_grad */
          double th0 = t[id] / m[id], thE = t[idE] / m[idE], thNE = t[idNE] / m[idNE], thN = t[idN] / m[idN];
          double the[3] = {interp(th0, thE), interp(th0, thNE), interp(th0, thN)};
          double gB[3] = {(w.bern[idE] - w.bern[id]) / LEN(0, id), (w.bern[id] - w.bern[idNE]) / LEN(1, id),
                          (w.bern[idN] - w.bern[id]) / LEN(2, id)};
          double gE[3] = {(w.exner[idE] - w.exner[id]) / LEN(0, id), (w.exner[id] - w.exner[idNE]) / LEN(1, id),
                          (w.exner[idN] - w.exner[id]) / LEN(2, id)};
          double phys[3] = {0.0, 0.0, 0.0};
          if (drt) { /* physics_velo_source: (-1)^(Lr-1) (visc_divu grad divu - visc_rotu curl rotu) */
            const double* D = ddv == 2 ? w.lapdivu : w.divu;
            double gd[3] = {0.0, 0.0, 0.0};
            if (ddv) {
              gd[0] = (D[idE] - D[id]) / LEN(0, id);
              gd[1] = (D[id] - D[idNE]) / LEN(1, id);
              gd[2] = (D[idN] - D[id]) / LEN(2, id);
            }
            const double cr[3] = {(w.vortL[id] - w.vortU[idS]) / PED(0, id), (w.vortL[id] - w.vortU[id]) / PED(1, id),
                                  (w.vortL[idW] - w.vortU[id]) / PED(2, id)};
            const double sg = (drt % 2) ? 1.0 : -1.0;
            for (int e = 0; e < 3; ++e) phys[e] = sg * (dif->visc_divu * gd[e] - dif->visc_rotu * cr[e]);
          }
          for (int e = 0; e < 3; ++e) {
            double src = -q[e] + phys[e] * LEN(e, id);
            du[e * NC + id] = src / LEN(e, id) - gB[e] - the[e] * gE[e];
            if (keep) keep->du_src[((int64_t)k * 3 + e) * NC + id] = src;
          }
        }
#undef HM
#undef QE
#undef U
#undef LEN
#undef PED
#undef AR
#undef UP_
#undef UD_
  }
  for (unsigned q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) free(*ptrs[q]);
  free(w.qe);
  free(w.hm);
  free(w.ht);
  free(w.ped_flux);
  if (dif) {
    free(w.len_flux); free(w.divu); free(w.lapdivu); free(w.lapm); free(w.lapt);
    free(w.vortL); free(w.vortU); free(w.rotu); free(gflux);
  }
}

static int is_master_cell(const wto_grid* g, int64_t c) {
  const int S = g->S, n = g->n;
  int64_t r = c % ((int64_t)S * S);
  int j = (int)(r / S) - H, i = (int)(r % S) - H;
  return (i >= 0 && i < n && j >= 0 && j < n) || c == g->south_pole || c == g->north_pole;
}

/* owned cells plus the pole masters: the cells an RK sub-step updates */
static int is_prognostic(const wto_grid* g, int64_t c) {
  const int S = g->S, n = g->n;
  int64_t r = c % ((int64_t)S * S);
  int j = (int)(r / S) - H, i = (int)(r % S) - H;
  return (i >= 0 && i < n && j >= 0 && j < n) || c == g->south_pole || c == g->north_pole;
}

static void rk4_impl(const wto_grid* g, const wto_diff* dif, int L, double dt, int nsteps, double* mass,
                     double* temp, double* velo);

void wto_rk4(const wto_grid* g, int L, double dt, int nsteps, double* mass, double* temp, double* velo) {
  rk4_impl(g, NULL, L, dt, nsteps, mass, temp, velo);
}

void wto_rk4_diff(const wto_grid* g, const wto_diff* dif, int L, double dt, int nsteps, double* mass, double* temp,
                  double* velo) {
  rk4_impl(g, dif, L, dt, nsteps, mass, temp, velo);
}

static void rk4_impl(const wto_grid* g, const wto_diff* dif, int L, double dt, int nsteps, double* mass,
                     double* temp, double* velo) {
  const int64_t NC = g->ncells;
  const int64_t ns = NC * L, ne = 3 * NC * L;
  double *m1 = malloc(sizeof(double) * ns), *t1 = malloc(sizeof(double) * ns), *u1 = malloc(sizeof(double) * ne);
  double *dm = malloc(sizeof(double) * ns), *dtm = malloc(sizeof(double) * ns), *du = malloc(sizeof(double) * ne);
  const double coef[4] = {dt / 4.0, dt / 3.0, dt / 2.0, dt};
  wto_ghost_fill(g, L, mass, temp, velo);
  for (int step = 0; step < nsteps; ++step) {
    memcpy(m1, mass, sizeof(double) * ns);
    memcpy(t1, temp, sizeof(double) * ns);
    memcpy(u1, velo, sizeof(double) * ne);
    for (int st = 0; st < 4; ++st) {
      const double* ym = st == 0 ? mass : m1;
      const double* yt = st == 0 ? temp : t1;
      const double* yu = st == 0 ? velo : u1;
      trend_impl(g, L, ym, yt, yu, dm, dtm, du, NULL, dif);
      double* om = st == 3 ? mass : m1;
      double* ot = st == 3 ? temp : t1;
      double* ou = st == 3 ? velo : u1;
      const double h = coef[st];
      WTO_PARALLEL_K_C
      for (int k = 0; k < L; ++k)
        for (int64_t c = 0; c < NC; ++c) {
          if (!is_prognostic(g, c)) continue;
          int64_t x = (int64_t)k * NC + c;
          om[x] = mass[x] + h * dm[x];
          ot[x] = temp[x] + h * dtm[x];
          for (int e = 0; e < 3; ++e) {
            int64_t y = ((int64_t)k * 3 + e) * NC + c;
            ou[y] = velo[y] + h * du[y];
          }
        }
      wto_ghost_fill(g, L, om, ot, ou);
    }
  }
  free(m1);
  free(t1);
  free(u1);
  free(dm);
  free(dtm);
  free(du);
}
