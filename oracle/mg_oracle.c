/* CPU oracle of the multi-level (wavelet) scalar transforms and the full
 * multigrid free-surface solver on a hierarchy of uniform levels of the
 * lozenge-tile layout.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * algorithms, used by tests/ as the checker.  Nothing in the product links
 * or calls it.
 *
 * Restated routines (reference src/):
 *   Compute_scalar_wavelets, Interp_node   wavelet.f90:648-670, 1487-1498
 *   Restrict_scalar                        wavelet.f90:520-564
 *   Prolong_scalar, Reconstruct_scalar     wavelet.f90:566-646 (Prolong_scalar's
 *                                          lifting has 11 terms: it omits the
 *                                          id2NE wavelet that Restrict_scalar uses)
 *   forward_scalar_transform (one level)   wavelet.f90:104-150
 *   inverse_scalar_transform (one level)   wavelet.f90:298-346
 *   FMG, v_cycle, Jacobi, prol_fun         lin_solve.f90:33-140, 233-299
 *   SRJ, SRJ_iter                          lin_solve.f90:142-231
 * Child cells of coarse cell (i, j) on the next finer level are (2i, 2j)
 * (coincident node) and the midpoints (2i+1, 2j) RT, (2i+1, 2j+1) DG,
 * (2i, 2j+1) UP of its edges, in the same lozenge frame.  Build with
 * -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mg_oracle.h"

static void ij_of(const wto_grid* g, int64_t c, int* i, int* j) {
  int64_t r = c % ((int64_t)g->S * g->S);
  *i = (int)(r % g->S) - 2;
  *j = (int)(r / g->S) - 2;
}

static int is_master(const wto_grid* g, int64_t c) {
  int i, j;
  ij_of(g, c, &i, &j);
  return (i >= 0 && i < g->n && j >= 0 && j < g->n) || c == g->south_pole || c == g->north_pole;
}

/* fill halo copies from owners */
static void ghost(const wto_grid* g, double* f) {
  for (int64_t c = 0; c < g->ncells; ++c) {
    int64_t s = g->node_src[c];
    if (s >= 0 && s != c) f[c] = f[s];
  }
}

/* Interp_node at fine midpoint cell x: its type from the parity of (i, j) */
static double interp_node(const wto_grid* gf, const double* ovl, const double* s, int64_t x) {
  const int64_t S = gf->S, NC = gf->ncells;
  int i, j;
  ij_of(gf, x, &i, &j);
  int64_t id1, id2, id3, id4;
  if ((i & 1) && !(j & 1)) { /* RT: Interp_node(idE, id, id2E, id2NE, id2S) */
    id1 = x - 1; id2 = x + 1; id3 = x + 1 + 2 * S; id4 = x - 1 - 2 * S;
  } else if ((i & 1) && (j & 1)) { /* DG: Interp_node(idNE, id2NE, id, id2E, id2N) */
    id1 = x + 1 + S; id2 = x - 1 - S; id3 = x + 1 - S; id4 = x - 1 + S;
  } else { /* UP: Interp_node(idN, id, id2N, id2W, id2NE) */
    id1 = x - S; id2 = x + S; id3 = x - 2 - S; id4 = x + 2 + S;
  }
  return (ovl[x] * s[id1] + ovl[NC + x] * s[id2] + ovl[2 * NC + x] * s[id3] + ovl[3 * NC + x] * s[id4]) *
         gf->hex_inv[x];
}

void wto_scalar_wavelets(const wto_grid* gf, const double* ovl, const double* s, double* wc) {
  for (int64_t x = 0; x < gf->ncells; ++x) {
    if (!is_master(gf, x)) continue;
    int i, j;
    ij_of(gf, x, &i, &j);
    if (!(i & 1) && !(j & 1)) continue; /* coincides with a coarse node */
    wc[x] = s[x] - interp_node(gf, ovl, s, x);
  }
  ghost(gf, wc);
}

/* lifting terms around fine cell c = 2p (wavelet.f90:550-562) */
static double lifting(const wto_grid* gf, const double* ovl, const double* wc, int64_t c, int with_2ne) {
  const int64_t S = gf->S, NC = gf->ncells;
  const double* a1 = ovl;
  const double* a2 = ovl + NC;
  const double* a3 = ovl + 2 * NC;
  const double* a4 = ovl + 3 * NC;
  const int64_t E = c + 1, NE = c + 1 + S, N2E = c + 2 + S, N2NE = c + 1 + 2 * S, N = c + S, W = c - 1,
                NW = c - 1 + S, S2W = c - 2 - S, SW = c - 1 - S, So = c - S, SW2 = c - 1 - 2 * S, SE = c + 1 - S;
  if (with_2ne)
    return wc[E] * a1[E] + wc[NE] * a2[NE] + wc[N2E] * a3[N2E] + wc[N2NE] * a4[N2NE] + wc[N] * a1[N] +
           wc[W] * a2[W] + wc[NW] * a3[NW] + wc[S2W] * a4[S2W] + wc[SW] * a1[SW] + wc[So] * a2[So] +
           wc[SW2] * a3[SW2] + wc[SE] * a4[SE];
  return wc[E] * a1[E] + wc[NE] * a2[NE] + wc[N2E] * a3[N2E] + wc[N] * a1[N] + wc[W] * a2[W] + wc[NW] * a3[NW] +
         wc[S2W] * a4[S2W] + wc[SW] * a1[SW] + wc[So] * a2[So] + wc[SW2] * a3[SW2] + wc[SE] * a4[SE];
}

/* Lifting at a pole (fine cell c is the pole master): the five edge
 * midpoints next to the pole, with the area they share with the pole's dual
 * cell (a2: the pole is their second end point), and the five DG midpoints of
 * the second ring, with the area on the pole's side (a3 at the south pole,
 * a4 at the north pole).  The reference's own pole stencil is its lozenge
 * stencil evaluated through the patch halo (wavelet.f90:542-567), where two
 * of the twelve neighbours fall on halo slots that no communication fills;
 * at the south pole the remaining terms are exactly these ten.  The inverse
 * (Prolong_scalar) drops the id2NE term, which at the poles is the DG
 * midpoint of lozenge 1 (south) / lozenge 7 (north). */
static double pole_lifting(const wto_grid* gf, const double* ovl, const double* wc, int64_t c, int prolong) {
  const int64_t S = gf->S, NC = gf->ncells, T = S * S;
  const int n = gf->n;
  double sum = 0.0;
  if (c == gf->south_pole) {
    for (int k = 0; k < 5; ++k) {
      const int64_t m = k * T + 2 * S + (n - 1 + 2), d = k * T + 3 * S + (n - 1 + 2);
      sum += wc[m] * ovl[NC + m];
      if (!(prolong && k == 1)) sum += wc[d] * ovl[2 * NC + d];
    }
  } else {
    for (int k = 5; k < 10; ++k) {
      const int64_t m = k * T + (n - 1 + 2) * S + 2, d = k * T + (n - 1 + 2) * S + 3;
      sum += wc[m] * ovl[NC + m];
      if (!(prolong && k == 7)) sum += wc[d] * ovl[3 * NC + d];
    }
  }
  return sum;
}

static int64_t child(const wto_grid* gf, const wto_grid* gc, int64_t p) {
  int i, j;
  ij_of(gc, p, &i, &j);
  int64_t lz = p / ((int64_t)gc->S * gc->S);
  return lz * gf->S * gf->S + (int64_t)(2 * j + 2) * gf->S + (2 * i + 2);
}

void wto_restrict(const wto_grid* gf, const wto_grid* gc, const double* ovl, const double* s_f, const double* wc,
                  double* s_c) {
  for (int64_t p = 0; p < gc->ncells; ++p) {
    if (!is_master(gc, p)) continue;
    const int64_t c = child(gf, gc, p);
    const int pole = p == gc->south_pole || p == gc->north_pole;
    s_c[p] = s_f[c] + (pole ? pole_lifting(gf, ovl, wc, c, 0) : lifting(gf, ovl, wc, c, 1)) * gc->hex_inv[p];
  }
  ghost(gc, s_c);
}

void wto_prolong(const wto_grid* gf, const wto_grid* gc, const double* ovl, const double* s_c, const double* wc,
                 double* s_f) {
  /* coarse-coincident fine nodes: undo the lifting */
  for (int64_t p = 0; p < gc->ncells; ++p) {
    if (!is_master(gc, p)) continue;
    const int64_t c = child(gf, gc, p);
    const int pole = p == gc->south_pole || p == gc->north_pole;
    s_f[c] = s_c[p] - (pole ? pole_lifting(gf, ovl, wc, c, 1) : lifting(gf, ovl, wc, c, 0)) * gc->hex_inv[p];
  }
  ghost(gf, s_f);
  /* midpoints: interpolate and add the wavelet */
  for (int64_t x = 0; x < gf->ncells; ++x) {
    if (!is_master(gf, x)) continue;
    int i, j;
    ij_of(gf, x, &i, &j);
    if (!(i & 1) && !(j & 1)) continue;
    s_f[x] = interp_node(gf, ovl, s_f, x) + wc[x];
  }
  ghost(gf, s_f);
}

/* ------------------------------------------------------------------------- */
/* FMG (lin_solve.f90:33-140) on levels 0..nl-1 (coarse to fine)              */
/* ------------------------------------------------------------------------- */
static void residual(const wto_mg* h, int l, const double* f, const double* u, double* res) {
  const wto_grid* g = h->grid[l];
  const int64_t NC = g->ncells;
  double* Au = (double*)malloc(NC * sizeof(double));
  wto_elliptic_apply(g, &h->p, h->eta_old[l], u, Au);
  for (int64_t c = 0; c < NC; ++c)
    if (is_master(g, c)) res[c] = f[c] - Au[c];
  ghost(g, res);
  free(Au);
}

static double l2(const wto_mg* h, int l, const double* a) { return sqrt(wto_wdot(h->grid[l], h->weight[l], a, a)); }

static double res_err(const wto_mg* h, int l, const double* f, const double* u, double nrm_f) {
  const int64_t NC = h->grid[l]->ncells;
  double* res = (double*)calloc(NC, sizeof(double));
  residual(h, l, f, u, res);
  double e = l2(h, l, res) / nrm_f;
  free(res);
  return e;
}

static void jacobi(const wto_mg* h, int l, double* u, const double* f, int n) {
  for (int i = 0; i < n; ++i) wto_jacobi(h->grid[l], &h->p, h->eta_old[l], u, f, 1.0);
}

/* prolongation with zero wavelets from level l-1 into level l of u */
static void prolong0(const wto_mg* h, int l, const double* uc, double* uf) {
  const int64_t NC = h->grid[l]->ncells;
  double* z = (double*)calloc(NC, sizeof(double));
  wto_prolong(h->grid[l], h->grid[l - 1], h->ovl[l], uc, z, uf);
  free(z);
}

/* restriction of level l of s into level l-1 (wavelets into a scratch) */
static void restrict1(const wto_mg* h, int l, const double* sf, double* sc) {
  const int64_t NC = h->grid[l]->ncells;
  double* wc = (double*)calloc(NC, sizeof(double));
  wto_scalar_wavelets(h->grid[l], h->ovl[l], sf, wc);
  wto_restrict(h->grid[l], h->grid[l - 1], h->ovl[l], sf, wc, sc);
  free(wc);
}

static void lc_add(const wto_grid* g, double* s, const double* a) {
  for (int64_t c = 0; c < g->ncells; ++c)
    if (is_master(g, c)) s[c] = 1.0 * s[c] + 1.0 * a[c];
  ghost(g, s);
}

void wto_fmg(const wto_mg* h, double** u, double* const* f, int* iters, double* res_out) {
  const int nl = h->nlevels;
  double nrm_f[32], nrm1[32];
  for (int l = 0; l < nl; ++l) {
    nrm_f[l] = l2(h, l, f[l]);
    if (nrm_f[l] == 0.0) nrm_f[l] = 1.0;
    nrm1[l] = res_err(h, l, f[l], u[l], nrm_f[l]);
    iters[l] = 0;
    res_out[l] = 0.0;
  }
  double err;
  iters[0] = wto_bicgstab(h->grid[0], &h->p, h->eta_old[0], h->weight[0], u[0], f[0], nrm_f[0], h->coarse_tol,
                          h->coarse_iter, &err);
  res_out[0] = err;
  double* corr[32];
  double* res[32];
  for (int l = 0; l < nl; ++l) {
    corr[l] = (double*)calloc(h->grid[l]->ncells, sizeof(double));
    res[l] = (double*)calloc(h->grid[l]->ncells, sizeof(double));
  }
  for (int l = 1; l < nl; ++l) {
    if (nrm1[l] < h->fine_tol) continue;
    prolong0(h, l, u[l - 1], u[l]);
    jacobi(h, l, u[l], f[l], h->pre_iter);
    for (int it = 1; it <= h->max_vcycle; ++it) {
      /* v_cycle */
      for (int j = 0; j <= l; ++j) memset(corr[j], 0, h->grid[j]->ncells * sizeof(double));
      residual(h, l, f[l], u[l], res[l]);
      for (int j = l; j >= 1; --j) {
        jacobi(h, j, corr[j], res[j], h->down_iter);
        residual(h, j, res[j], corr[j], res[j]);
        restrict1(h, j, res[j], res[j - 1]);
      }
      double nrm_rhs = l2(h, 0, res[0]);
      if (nrm_rhs == 0.0) nrm_rhs = 1.0;  /* tol = 0 in the reference's comparison (lin_solve.f90:124) */
      wto_bicgstab(h->grid[0], &h->p, h->eta_old[0], h->weight[0], corr[0], res[0], nrm_rhs, h->coarse_tol,
                   h->coarse_iter, &err);
      for (int j = 1; j <= l; ++j) {
        double* pf = (double*)malloc(h->grid[j]->ncells * sizeof(double));
        memcpy(pf, corr[j], h->grid[j]->ncells * sizeof(double));
        prolong0(h, j, corr[j - 1], pf);
        lc_add(h->grid[j], corr[j], pf);
        free(pf);
        jacobi(h, j, corr[j], res[j], h->up_iter);
      }
      lc_add(h->grid[l], u[l], corr[l]);
      jacobi(h, l, u[l], f[l], h->post_iter);
      res_out[l] = res_err(h, l, f[l], u[l], nrm_f[l]);
      iters[l] = it;
      if (res_out[l] <= h->fine_tol) break;
    }
  }
  for (int l = 0; l < nl; ++l) {
    free(corr[l]);
    free(res[l]);
  }
}

/* ------------------------------------------------------------------------- */
/* SRJ (lin_solve.f90:142-231): BiCGSTAB at the coarsest level, then on each  */
/* finer level a zero-wavelet prolongation followed by scheduled-relaxation   */
/* Jacobi sweeps, weights w(n) = 2 / (k1 + k2 - (k2 - k1) cos(pi (2n-1)/2m)). */
/* ------------------------------------------------------------------------- */
void wto_srj(const wto_mg* h, int max_srj_iter, double** u, double* const* f, int* iters, double* res_out) {
  enum { M = 8 };
  const double k1 = 3e-2, k2 = 1.8, pi = acos(-1.0);
  const int nl = h->nlevels;
  double w[M], nrm_f[32];
  if (max_srj_iter < M) {
    for (int i = 0; i < M; ++i) w[i] = 1.0;
  } else {
    max_srj_iter = M * (max_srj_iter / M);
    for (int i = 1; i <= M; ++i) w[i - 1] = 2.0 / (k1 + k2 - (k2 - k1) * cos(pi * (2.0 * (double)i - 1.0) / (2.0 * (double)M)));
  }
  for (int l = 0; l < nl; ++l) {
    nrm_f[l] = l2(h, l, f[l]);
    if (nrm_f[l] == 0.0) nrm_f[l] = 1.0;
    iters[l] = 0;
    res_out[l] = 0.0;
  }
  double err;
  iters[0] = wto_bicgstab(h->grid[0], &h->p, h->eta_old[0], h->weight[0], u[0], f[0], nrm_f[0], h->coarse_tol,
                          h->coarse_iter, &err);
  res_out[0] = err;
  for (int l = 1; l < nl; ++l) {
    prolong0(h, l, u[l - 1], u[l]);
    int ii = 0;
    for (int it = 1; it <= max_srj_iter; ++it) {
      if (++ii > M) {
        if (res_out[l] < h->fine_tol) break;
        ii = 1;
      }
      wto_jacobi(h->grid[l], &h->p, h->eta_old[l], u[l], f[l], w[ii - 1]);
      res_out[l] = res_err(h, l, f[l], u[l], nrm_f[l]);
      iters[l] = it;
      if (res_out[l] <= h->fine_tol) break;
    }
  }
}

