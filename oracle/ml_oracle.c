/* CPU oracle of the multi-level explicit step on a full hierarchy of uniform
 * levels (every mask TOLRNZ, as the reference's uniform refinement with zero
 * thresholds leaves them), in the lozenge-tile layout.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * algorithms, used by tests/ as the checker.  Nothing in the product links
 * or calls it.
 *
 * Restated routines (reference src/):
 *   Compute_velo_wavelets, Interp_outer_velo,       wavelet.f90:672-707, 1054-1094,
 *   Interp_inner_velo                               1000-1052
 *   Compute_velo_wavelets_penta,                    wavelet.f90:709-781, 897-925
 *   velo_interp_penta_corr
 *   Reconstruct_outer_velo, Reconstruct_inner_velo  wavelet.f90:954-998
 *   Reconstruct_velo_penta                          wavelet.f90:1096-1152
 *   trend_ml over levels with compute-or-restrict   multi_level.f90:5-188, 311-577
 *     (scalar_cpt_restr, flux_cpt_restr, interp_flux, complete_coarse_flux,
 *      part_coarse_flux, coarse_flux, get_indices, u_source_cpt_restr)
 *   WT_after_step                                   adapt.f90:192-237
 *   inverse_wavelet_transform                       wavelet.f90:208-296
 *   RK4                                             time_integr.f90:78-103
 * On a full hierarchy every coarse node has children, so the coarse
 * Bernoulli and Exner functions, fluxes and velocity sources are all
 * restricted from the next finer level; the coarse step1 values they
 * replace are not computed.  Child cells of coarse cell (i, j) are (2i, 2j)
 * and the midpoints (2i+1, 2j) RT, (2i+1, 2j+1) DG, (2i, 2j+1) UP in the
 * same lozenge frame.  All four lozenge corners are pentagons.  Build with
 * -ffp-contract=off.
 */
#include "ml_oracle.h"

#include <stdlib.h>
#include <string.h>

#include "elliptic_oracle.h"
#include "mg_oracle.h"
#include "split_oracle.h"

#define H 2

static inline int64_t cell_of(const wto_grid* g, int lz, int i, int j) {
  return (int64_t)lz * g->S * g->S + (int64_t)(j + H) * g->S + (int64_t)(i + H);
}

/* shared.f90:282-330 tables, 0-based edge index e = RT 0, DG 1, UP 2 */
static const int HSO[3] = {2, 0, 4};
static const int HEXS[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {-1, 0, 0}, {-1, -1, 1},
                                {0, -1, 2}, {0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {-1, 0, 0}};
static const int NGHB[10][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 0}};
static const int ENDPT[3][2][2] = {{{0, 0}, {1, 0}}, {{1, 1}, {0, 0}}, {{0, 0}, {0, 1}}};
static const int OPPNO[3][2][2] = {{{0, -1}, {1, 1}}, {{0, 1}, {1, 0}}, {{1, 1}, {-1, 0}}};
static const int BFLY2[3][4][2] = {{{-3, -2}, {1, -2}, {3, 2}, {-1, 2}},
                                   {{1, 3}, {-3, -1}, {-1, -3}, {3, 1}},
                                   {{2, -1}, {2, 3}, {-2, 1}, {-2, -3}}};
/* Iu_Base_Wgt (wavelet.f90:1765) */
static const double BASE[9] = {16.0 / 16, -1.0 / 16, 1.0 / 16, 1.0 / 16, -1.0 / 16, -1.0 / 16, -1.0 / 16, 1.0 / 16,
                               1.0 / 16};

static void edge_ghost(const wto_grid* g, double* u) {
  const int64_t NC = g->ncells;
  for (int64_t c = 0; c < NC; ++c)
    for (int e = 0; e < 3; ++e) {
      const int64_t slot = 3 * c + e, s = g->edge_src[slot];
      if (s >= 0 && s != slot) u[e * NC + c] = (double)g->edge_sign[slot] * u[(s % 3) * NC + s / 3];
    }
}

/* ---------------------------------------------------------------------------
 * velocity transform
 * ------------------------------------------------------------------------- */
/* Interp_outer_velo: coarse edge e of parent (ip, jp), weights at fine cell id2 */
static double interp_outer(const wto_ml* h, int l, int lz, int ip, int jp, int e, int64_t id2, const double* uc) {
  const wto_grid* gc = h->grid[l - 1];
  const int64_t NCc = gc->ncells, NCf = h->grid[l]->ncells;
  const int so = HSO[e];
  const int* e1 = ENDPT[e][0];
  const int* e2 = ENDPT[e][1];
  const int* o1 = OPPNO[e][0];
  const int* o2 = OPPNO[e][1];
#define VC(o, col) uc[(int64_t)HEXS[col][2] * NCc + cell_of(gc, lz, ip + (o)[0] + HEXS[col][0], jp + (o)[1] + HEXS[col][1])]
  double v[9];
  v[0] = uc[(int64_t)e * NCc + cell_of(gc, lz, ip, jp)];
  v[1] = VC(e2, so + 2);
  v[2] = VC(e1, so + 3);
  v[3] = VC(e1, so + 5);
  v[4] = VC(e2, so + 0);
  v[5] = VC(o1, so + 1) - VC(e1, so + 2);
  v[6] = VC(e2, so + 3) - VC(o1, so + 4);
  v[7] = VC(o2, so + 4) - VC(e2, so + 5);
  v[8] = VC(e1, so + 0) - VC(o2, so + 1);
#undef VC
  const double* iu = h->iu[l];
  double s = 0.0;
  for (int m = 0; m < 9; ++m) {
    const double w = BASE[m] + iu[(int64_t)m * NCf + id2];
    s = s + w * v[m];
  }
  return s;
}

/* Interp_inner_velo: the six inner child edges of parent (ip, jp) */
static void interp_inner(const wto_ml* h, int l, int lz, int ip, int jp, const double* uc, const double* uf,
                         double out[6]) {
  const wto_grid* gc = h->grid[l - 1];
  const wto_grid* gf = h->grid[l];
  const int64_t NCc = gc->ncells, NCf = gf->ncells;
  const int64_t idp = cell_of(gc, lz, ip, jp);
  double curl[2];
  for (int t = 0; t < 2; ++t) {
    const int64_t id1 = cell_of(gc, lz, ip - t + 1, jp), id2 = cell_of(gc, lz, ip, jp + t);
    curl[t] = (uc[NCc + idp] * gc->len[NCc + idp] + uc[2 * NCc + id1] * gc->len[2 * NCc + id1] +
               uc[id2] * gc->len[id2]) /
              gc->triarea[(int64_t)t * NCc + idp];
  }
  const int ic = 2 * ip, jc = 2 * jp;
  const int64_t id = cell_of(gf, lz, ic, jc), idE = cell_of(gf, lz, ic + 1, jc), idNE = cell_of(gf, lz, ic + 1, jc + 1),
                idN = cell_of(gf, lz, ic, jc + 1), idN2E = cell_of(gf, lz, ic + 2, jc + 1),
                id2NE = cell_of(gf, lz, ic + 1, jc + 2), idN2 = cell_of(gf, lz, ic, jc + 2),
                idE2 = cell_of(gf, lz, ic + 2, jc);
  const double* len = gf->len;
  const double* tri = gf->triarea;
#define UF(e, c) (uf[(int64_t)(e) * NCf + (c)])
#define LF(e, c) (len[(int64_t)(e) * NCf + (c)])
  out[0] = (tri[id] * curl[0] - UF(0, id) * LF(0, id) - UF(1, id) * LF(1, id)) / LF(2, idE);
  out[1] = (tri[idE] * curl[0] - UF(0, idE) * LF(0, idE) - UF(2, idE2) * LF(2, idE2)) / LF(1, idE);
  out[2] = (tri[idNE] * curl[0] - UF(1, idNE) * LF(1, idNE) - UF(2, idN2E) * LF(2, idN2E)) / LF(0, idNE);
  out[3] = (tri[NCf + id] * curl[1] - UF(2, id) * LF(2, id) - UF(1, id) * LF(1, id)) / LF(0, idN);
  out[4] = (tri[NCf + idN] * curl[1] - UF(2, idN) * LF(2, idN) - UF(0, idN2) * LF(0, idN2)) / LF(1, idN);
  out[5] = (tri[NCf + idNE] * curl[1] - UF(1, idNE) * LF(1, idNE) - UF(0, id2NE) * LF(0, id2NE)) / LF(2, idNE);
#undef UF
#undef LF
}

/* inner child edges in the order of Interp_inner_velo: (cell, edge) */
static void inner_slots(const wto_grid* gf, int lz, int ic, int jc, int64_t slot[6][2]) {
  const int64_t idE = cell_of(gf, lz, ic + 1, jc), idNE = cell_of(gf, lz, ic + 1, jc + 1), idN = cell_of(gf, lz, ic, jc + 1);
  slot[0][0] = idE;  slot[0][1] = 2;
  slot[1][0] = idE;  slot[1][1] = 1;
  slot[2][0] = idNE; slot[2][1] = 0;
  slot[3][0] = idN;  slot[3][1] = 0;
  slot[4][0] = idN;  slot[4][1] = 1;
  slot[5][0] = idNE; slot[5][1] = 2;
}

/* pentagon corrections (Compute_velo_wavelets_penta sign = +1 on wavelets,
 * Reconstruct_velo_penta sign = -1 on velocities): the corrections that the
 * two routines add, expressed as the change of wc; velocities change by the
 * opposite amount */
static void penta(const wto_ml* h, int l, int lz, const double* uc, double* f, int sign) {
  const wto_grid* gc = h->grid[l - 1];
  const wto_grid* gf = h->grid[l];
  const int64_t NCc = gc->ncells, NCf = gf->ncells;
  const int nc = gc->n, nf = gf->n;
  const double* iu = h->iu[l];
#define UC(e, i, j) uc[(int64_t)(e) * NCc + cell_of(gc, lz, i, j)]
#define F(e, c) f[(int64_t)(e) * NCf + (c)]
  { /* IMINUSJPLUS (upper-left corner) */
    const int64_t id = cell_of(gf, lz, 0, nf - 2), idN = cell_of(gf, lz, 0, nf - 1);
    double v = (BASE[7] + iu[7 * NCf + idN]) * (UC(2, 0, nc) + UC(0, -1, nc));
    if (sign > 0) v = -v;                      /* Compute: v = -(w)(...) ; wc(id) -= v */
    if (sign > 0) {
      F(2, id) = F(2, id) - v;
      F(2, idN) = F(2, idN) + v;
    } else {
      F(2, id) = F(2, id) - v;
      F(2, idN) = F(2, idN) + v;
    }
  }
  { /* IPLUSJMINUS (lower-right corner) */
    const int64_t id = cell_of(gf, lz, nf - 2, 0), idE = cell_of(gf, lz, nf - 1, 0);
    double v = (BASE[6] + iu[6 * NCf + idE]) * (UC(0, nc, 0) + UC(2, nc, -1));
    if (sign < 0) v = -v;                      /* Reconstruct: v = -(w)(...) */
    F(0, id) = F(0, id) - v;
    F(0, idE) = F(0, idE) + v;
  }
  { /* IJMINUS (lower-left corner): velo_interp_penta_corr */
    const int64_t id = cell_of(gf, lz, 0, 0), idN = cell_of(gf, lz, 0, 1), idE = cell_of(gf, lz, 1, 0);
    const double u1 = (BASE[8] + iu[8 * NCf + idN]) *
                      ((-UC(2, 0, -1) - (-UC(0, -1, -1))) - (UC(1, -1, -1) - UC(2, -1, -1)));
    const double u2 = (BASE[5] + iu[5 * NCf + idE]) * (UC(0, -1, -1) + UC(0, -1, 0) - (UC(0, -1, -1) - UC(1, -1, -1)));
    if (sign > 0) {
      F(2, id) = F(2, id) + u1;
      F(2, idN) = F(2, idN) - u1;
      F(0, id) = F(0, id) + u2;
      F(0, idE) = F(0, idE) - u2;
    } else {
      F(2, id) = F(2, id) - u1;
      F(2, idN) = F(2, idN) + u1;
      F(0, id) = F(0, id) - u2;
      F(0, idE) = F(0, idE) + u2;
    }
  }
#undef UC
#undef F
}

void wto_velo_wavelets(const wto_ml* h, int l, const double* u_f, const double* u_c, double* wc) {
  const wto_grid* gc = h->grid[l - 1];
  const wto_grid* gf = h->grid[l];
  const int64_t NCf = gf->ncells;
  memset(wc, 0, sizeof(double) * 3 * NCf);
  for (int lz = 0; lz < 10; ++lz) {
    for (int jp = 0; jp < gc->n; ++jp)
      for (int ip = 0; ip < gc->n; ++ip) {
        const int ic = 2 * ip, jc = 2 * jp;
        for (int e = 0; e < 3; ++e) {
          const int64_t id1 = cell_of(gf, lz, ic + ENDPT[e][0][0], jc + ENDPT[e][0][1]);
          const int64_t id2 = cell_of(gf, lz, ic + ENDPT[e][1][0], jc + ENDPT[e][1][1]);
          const double u = interp_outer(h, l, lz, ip, jp, e, id2, u_c);
          wc[e * NCf + id2] = u_f[e * NCf + id2] - u;
          wc[e * NCf + id1] = -wc[e * NCf + id2];
        }
        double ui[6];
        int64_t sl[6][2];
        interp_inner(h, l, lz, ip, jp, u_c, u_f, ui);
        inner_slots(gf, lz, ic, jc, sl);
        for (int k = 0; k < 6; ++k) {
          const int64_t x = sl[k][1] * NCf + sl[k][0];
          wc[x] = u_f[x] - ui[k];
        }
      }
    penta(h, l, lz, u_c, wc, +1);
  }
  edge_ghost(gf, wc);
}

void wto_velo_reconstruct(const wto_ml* h, int l, const double* u_c, const double* wc, double* u_f) {
  const wto_grid* gc = h->grid[l - 1];
  const wto_grid* gf = h->grid[l];
  const int64_t NCf = gf->ncells, NCc = gc->ncells;
  for (int lz = 0; lz < 10; ++lz) {
    for (int jp = 0; jp < gc->n; ++jp)
      for (int ip = 0; ip < gc->n; ++ip) {
        const int ic = 2 * ip, jc = 2 * jp;
        const int64_t idp = cell_of(gc, lz, ip, jp);
        for (int e = 0; e < 3; ++e) {
          const int64_t id1 = cell_of(gf, lz, ic + ENDPT[e][0][0], jc + ENDPT[e][0][1]);
          const int64_t id2 = cell_of(gf, lz, ic + ENDPT[e][1][0], jc + ENDPT[e][1][1]);
          const double u = interp_outer(h, l, lz, ip, jp, e, id2, u_c);
          u_f[e * NCf + id2] = u + wc[e * NCf + id2];
          u_f[e * NCf + id1] = 2.0 * u_c[e * NCc + idp] - u_f[e * NCf + id2];
        }
      }
    penta(h, l, lz, u_c, u_f, -1);
  }
  edge_ghost(gf, u_f);
  for (int lz = 0; lz < 10; ++lz)
    for (int jp = 0; jp < gc->n; ++jp)
      for (int ip = 0; ip < gc->n; ++ip) {
        double ui[6];
        int64_t sl[6][2];
        interp_inner(h, l, lz, ip, jp, u_c, u_f, ui);
        inner_slots(gf, lz, 2 * ip, 2 * jp, sl);
        for (int k = 0; k < 6; ++k) {
          const int64_t x = sl[k][1] * NCf + sl[k][0];
          u_f[x] = ui[k] + wc[x];
        }
      }
  edge_ghost(gf, u_f);
}

/* ---------------------------------------------------------------------------
 * flux restriction
 * ------------------------------------------------------------------------- */
enum { UZM = 0, UPZ, UMZ, UZP, VMM, VPM, VMP, VPP, WMM, WPM, WMP, WPP, WMMM, WPPP, VMPP, VPMM };

typedef struct {
  int64_t ed[16]; /* edge plane offset e*NC + cell */
  int64_t mp, pp, pm, mm;
} indices_t;

/* get_indices (multi_level.f90:619-659) at fine (i, j), edge e */
static void get_indices(const wto_grid* g, int lz, int i, int j, int e, indices_t* x) {
  const int so = HSO[e];
  const int64_t NC = g->ncells;
#define ED(bi, bj, col) ((int64_t)HEXS[col][2] * NC + cell_of(g, lz, (bi) + HEXS[col][0], (bj) + HEXS[col][1]))
  x->ed[UMZ] = ED(i, j, so + 1);
  x->ed[UPZ] = ED(i, j, so + 4);
  x->ed[WMP] = ED(i, j, so + 0);
  x->ed[VPP] = ED(i, j, so + 5);
  x->ed[WPM] = ED(i, j, so + 3);
  x->ed[VMM] = ED(i, j, so + 2);
  const int mpi = i + NGHB[so + 0][0], mpj = j + NGHB[so + 0][1];
  const int ppi = i + NGHB[so + 5][0], ppj = j + NGHB[so + 5][1];
  const int pmi = i + NGHB[so + 3][0], pmj = j + NGHB[so + 3][1];
  const int mmi = i + NGHB[so + 2][0], mmj = j + NGHB[so + 2][1];
  x->mp = cell_of(g, lz, mpi, mpj);
  x->pp = cell_of(g, lz, ppi, ppj);
  x->pm = cell_of(g, lz, pmi, pmj);
  x->mm = cell_of(g, lz, mmi, mmj);
  x->ed[VMP] = ED(mpi, mpj, so + 2);
  x->ed[VMPP] = ED(mpi, mpj, so + 5);
  x->ed[UZP] = ED(mpi, mpj, so + 4);
  x->ed[WPPP] = ED(ppi, ppj, so + 0);
  x->ed[WPP] = ED(ppi, ppj, so + 3);
  x->ed[VPM] = ED(pmi, pmj, so + 5);
  x->ed[VPMM] = ED(pmi, pmj, so + 2);
  x->ed[UZM] = ED(pmi, pmj, so + 1);
  x->ed[WMMM] = ED(mmi, mmj, so + 3);
  x->ed[WMM] = ED(mmi, mmj, so + 0);
#undef ED
}

/* scalar trend at a fine lattice point, through the wide map beyond the tile */
static double dval(const wto_ml* h, int l, int lz, int i, int j, const double* df) {
  const wto_grid* g = h->grid[l];
  if (i >= -H && j >= -H && i < g->n + H && j < g->n + H) return df[cell_of(g, lz, i, j)];
  const int WH = h->WH[l], WS = g->n + 2 * WH;
  const int64_t s = h->wide_src[l][(int64_t)lz * WS * WS + (int64_t)(j + WH) * WS + (i + WH)];
  return s >= 0 ? df[s] : 0.0;
}

static double sum3(const double* hf, const int64_t* ed, const int a[3], const double* rf, int64_t NCc, int64_t p) {
  double s = 0.0;
  for (int m = 0; m < 3; ++m) s = s + hf[ed[a[m]]] * rf[(int64_t)m * NCc + p];
  return s;
}
static double sum3d(const double* hf, const int64_t* ed, const int a[3], const int b[3], const double* rf, int64_t NCc,
                    int64_t p) {
  double s = 0.0;
  for (int m = 0; m < 3; ++m) s = s + (hf[ed[a[m]]] - hf[ed[b[m]]]) * rf[(int64_t)m * NCc + p];
  return s;
}

/* part_coarse_flux (multi_level.f90:504-535) at fine (i, j), edge e */
static double part_coarse_flux(const wto_ml* h, int l, int lz, int i, int j, int e, const double* hf, const double* df) {
  const wto_grid* g = h->grid[l];
  const int64_t NC = g->ncells;
  const double* a = h->ovl[l];
  const double* sp = h->split[l];
  indices_t x;
  get_indices(g, lz, i, j, e, &x);
  const int64_t c = cell_of(g, lz, i, j);
  double area[2] = {a[c], a[NC + c]};
  double ol[4];
  ol[0] = sp[c];
  ol[1] = sp[NC + c];
  ol[2] = a[2 * NC + c] - ol[0];
  ol[3] = a[3 * NC + c] - ol[1];
  area[0] = area[0] + ol[0] + ol[3];
  area[1] = area[1] + ol[1] + ol[2];
  const double tot = area[0] + area[1];
  area[0] = area[0] / tot;
  area[1] = area[1] / tot;
  ol[0] = sp[x.pp];
  ol[1] = sp[NC + x.mm];
  ol[2] = a[2 * NC + x.mp] - sp[x.mp];
  ol[3] = a[3 * NC + x.pm] - sp[NC + x.pm];
  const double s1 = hf[x.ed[UPZ]] * area[0] + hf[x.ed[UMZ]] * area[1];
  const double s2 = hf[x.ed[VMM]] + hf[x.ed[WMP]];
  const double s3 = hf[x.ed[WPM]] + hf[x.ed[VPP]];
  return s1 - s2 * area[1] - s3 * area[0] + ol[2] * df[x.mp] - ol[3] * df[x.pm] - ol[0] * df[x.pp] +
         ol[1] * df[x.mm];
}

/* coarse_flux (multi_level.f90:537-577) at fine (i, j), edge e */
static double coarse_flux(const wto_ml* h, int l, int lz, int i, int j, int e, const double* df) {
  const wto_grid* g = h->grid[l];
  const int64_t NC = g->ncells;
  const double* a = h->ovl[l];
  const double* hi = g->hex_inv;
  const int so = HSO[e];
#define NB(k) cell_of(g, lz, i + NGHB[k][0], j + NGHB[k][1])
  const int64_t mz = NB(so + 1), pz = NB(so + 4), mp = NB(so + 0), pp = NB(so + 5), pm = NB(so + 3), mm = NB(so + 2);
#undef NB
  const double dmm2 = dval(h, l, lz, i + BFLY2[e][0][0], j + BFLY2[e][0][1], df);
  const double dpm2 = dval(h, l, lz, i + BFLY2[e][1][0], j + BFLY2[e][1][1], df);
  const double dpp2 = dval(h, l, lz, i + BFLY2[e][2][0], j + BFLY2[e][2][1], df);
  const double dmp2 = dval(h, l, lz, i + BFLY2[e][3][0], j + BFLY2[e][3][1], df);
  const int64_t id = cell_of(g, lz, i, j);
#define A(k, c) a[(int64_t)((k) - 1) * NC + (c)]
  return (A(1, id) * A(2, id) * hi[id] + A(2, mp) * A(3, mp) * hi[mp] + A(1, pp) * A(3, pp) * hi[pp] +
          A(1, pm) * A(4, pm) * hi[pm] + A(2, mm) * A(4, mm) * hi[mm]) *
             (df[pz] - df[mz]) +
         A(3, pp) * A(4, pp) * hi[pp] * 0.5 * (dpp2 - df[mz]) + A(3, pm) * A(4, pm) * hi[pm] * 0.5 * (dpm2 - df[mz]) +
         A(3, mp) * A(4, mp) * hi[mp] * 0.5 * (df[pz] - dmp2) + A(3, mm) * A(4, mm) * hi[mm] * 0.5 * (df[pz] - dmm2);
#undef A
}

void wto_flux_restrict(const wto_ml* h, int l, const double* hf, const double* df, double* hc) {
  const wto_grid* gc = h->grid[l - 1];
  const wto_grid* gf = h->grid[l];
  const int64_t NCc = gc->ncells;
  const double* rf = h->rf[l - 1];
  static const int A0[3] = {WPM, UZM, VMM}, B0a[3] = {VPM, WMMM, UMZ}, B0b[3] = {UPZ, VPMM, WMM};
  static const int A1[3] = {WMP, UZP, VPP}, B1a[3] = {VMP, WPPP, UPZ}, B1b[3] = {UMZ, VMPP, WPP};
  static const int A2[3] = {UZM, VMM, WPM}, B2a[3] = {WMMM, UMZ, VPM}, B2b[3] = {VPMM, WMM, UPZ};
  static const int A3[3] = {UZP, VPP, WMP}, B3a[3] = {WPPP, UPZ, VMP}, B3b[3] = {VMPP, WPP, UMZ};
  memset(hc, 0, sizeof(double) * 3 * NCc);
  for (int lz = 0; lz < 10; ++lz)
    for (int jp = 0; jp < gc->n; ++jp)
      for (int ip = 0; ip < gc->n; ++ip) {
        const int i = 2 * ip, j = 2 * jp;
        const int64_t p = cell_of(gc, lz, ip, jp);
        indices_t x;
        double sm[4];
        get_indices(gf, lz, i + 1, j, 0, &x);
        sm[0] = -sum3(hf, x.ed, A0, rf + 0 * NCc, NCc, p) - sum3d(hf, x.ed, B0a, B0b, rf + 3 * NCc, NCc, p);
        sm[1] = sum3(hf, x.ed, A1, rf + 6 * NCc, NCc, p) + sum3d(hf, x.ed, B1a, B1b, rf + 9 * NCc, NCc, p);
        get_indices(gf, lz, i, j + 1, 2, &x);
        sm[2] = -sum3(hf, x.ed, A2, rf + 12 * NCc, NCc, p) - sum3d(hf, x.ed, B2a, B2b, rf + 15 * NCc, NCc, p);
        sm[3] = sum3(hf, x.ed, A3, rf + 18 * NCc, NCc, p) + sum3d(hf, x.ed, B3a, B3b, rf + 21 * NCc, NCc, p);
        double pf, cf;
        pf = part_coarse_flux(h, l, lz, i + 1, j, 0, hf, df);
        cf = coarse_flux(h, l, lz, i + 1, j, 0, df);
        hc[p] = pf + cf + sm[0] + sm[1];
        pf = part_coarse_flux(h, l, lz, i + 1, j + 1, 1, hf, df);
        cf = coarse_flux(h, l, lz, i + 1, j + 1, 1, df);
        hc[NCc + p] = pf + cf + sm[1] + sm[2];
        pf = part_coarse_flux(h, l, lz, i, j + 1, 2, hf, df);
        cf = coarse_flux(h, l, lz, i, j + 1, 2, df);
        hc[2 * NCc + p] = pf + cf + sm[2] + sm[3];
      }
  edge_ghost(gc, hc);
}

/* ---------------------------------------------------------------------------
 * trend on the hierarchy
 * ------------------------------------------------------------------------- */
static int prognostic(const wto_grid* g, int64_t c) {
  int64_t r = c % ((int64_t)g->S * g->S);
  int j = (int)(r / g->S) - H, i = (int)(r % g->S) - H;
  return (i >= 0 && i < g->n && j >= 0 && j < g->n) || c == g->south_pole || c == g->north_pole;
}

static void divergence(const wto_grid* g, const double* hf, double* d) {
  const int64_t NC = g->ncells;
  const int n = g->n;
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i) {
        const int64_t c = cell_of(g, lz, i, j);
        if (!prognostic(g, c)) continue;
        const int64_t cW = cell_of(g, lz, i - 1, j), cS = cell_of(g, lz, i, j - 1), cSW = cell_of(g, lz, i - 1, j - 1);
        const double div = (hf[c] - hf[cW] + hf[NC + cSW] - hf[NC + c] + hf[2 * NC + c] - hf[2 * NC + cS]) * g->hex_inv[c];
        d[c] = -div;
      }
}

void wto_trend_ml(const wto_ml* h, int L, double* const* mass, double* const* temp, double* const* velo,
                  double** dmass, double** dtemp, double** dvelo) {
  const int nl = h->nlevels, top = nl - 1;
  const wto_grid* gt = h->grid[top];
  const int64_t NT = gt->ncells;
  wto_trend_keep keep;
  keep.hflux_m = malloc(sizeof(double) * 3 * NT * L);
  keep.hflux_t = malloc(sizeof(double) * 3 * NT * L);
  keep.bern = malloc(sizeof(double) * NT * L);
  keep.exner = malloc(sizeof(double) * NT * L);
  keep.du_src = calloc((size_t)3 * NT * L, sizeof(double));
  wto_trend_parts(gt, L, mass[top], temp[top], velo[top], dmass[top], dtemp[top], dvelo[top], &keep);
  /* per-level work of the layer being restricted */
  double *hm[16], *ht[16], *bern[16], *exner[16], *src[16];
  for (int l = 0; l < top; ++l) {
    const int64_t NC = h->grid[l]->ncells;
    hm[l] = malloc(sizeof(double) * 3 * NC);
    ht[l] = malloc(sizeof(double) * 3 * NC);
    bern[l] = calloc(NC, sizeof(double));
    exner[l] = calloc(NC, sizeof(double));
    src[l] = calloc(3 * NC, sizeof(double));
    memset(dmass[l], 0, sizeof(double) * NC * L);
    memset(dtemp[l], 0, sizeof(double) * NC * L);
    memset(dvelo[l], 0, sizeof(double) * 3 * NC * L);
  }
  for (int k = 0; k < L; ++k) {
    hm[top] = keep.hflux_m + (int64_t)k * 3 * NT;
    ht[top] = keep.hflux_t + (int64_t)k * 3 * NT;
    bern[top] = keep.bern + (int64_t)k * NT;
    exner[top] = keep.exner + (int64_t)k * NT;
    src[top] = keep.du_src + (int64_t)k * 3 * NT;
    /* the finest fluxes and scalar trends are ghost filled before restriction */
    edge_ghost(gt, hm[top]);
    edge_ghost(gt, ht[top]);
    wto_node_ghost(gt, dmass[top] + (int64_t)k * NT);
    wto_node_ghost(gt, dtemp[top] + (int64_t)k * NT);
    for (int l = top - 1; l >= 0; --l) {
      const wto_grid* gc = h->grid[l];
      const wto_grid* gf = h->grid[l + 1];
      const int64_t NC = gc->ncells, NF = gf->ncells;
      /* Bernoulli and Exner by sub-sampling (scalar_cpt_restr) on [0, n]^2 */
      for (int lz = 0; lz < 10; ++lz)
        for (int j = 0; j <= gc->n; ++j)
          for (int i = 0; i <= gc->n; ++i) {
            const int64_t c = cell_of(gc, lz, i, j), f = cell_of(gf, lz, 2 * i, 2 * j);
            bern[l][c] = bern[l + 1][f];
            exner[l][c] = exner[l + 1][f];
          }
      double* dm = dmass[l] + (int64_t)k * NC;
      double* dt = dtemp[l] + (int64_t)k * NC;
      wto_flux_restrict(h, l + 1, hm[l + 1], dmass[l + 1] + (int64_t)k * NF, hm[l]);
      wto_flux_restrict(h, l + 1, ht[l + 1], dtemp[l + 1] + (int64_t)k * NF, ht[l]);
      divergence(gc, hm[l], dm);
      divergence(gc, ht[l], dt);
      wto_node_ghost(gc, dm);
      wto_node_ghost(gc, dt);
      /* velocity source: sum of the two child edges (u_source_cpt_restr) */
      for (int lz = 0; lz < 10; ++lz)
        for (int jp = 0; jp < gc->n; ++jp)
          for (int ip = 0; ip < gc->n; ++ip) {
            const int64_t p = cell_of(gc, lz, ip, jp);
            const int64_t id = cell_of(gf, lz, 2 * ip, 2 * jp), idE = cell_of(gf, lz, 2 * ip + 1, 2 * jp),
                          idNE = cell_of(gf, lz, 2 * ip + 1, 2 * jp + 1), idN = cell_of(gf, lz, 2 * ip, 2 * jp + 1);
            src[l][p] = src[l + 1][id] + src[l + 1][idE];
            src[l][NC + p] = src[l + 1][NF + id] + src[l + 1][NF + idNE];
            src[l][2 * NC + p] = src[l + 1][2 * NF + id] + src[l + 1][2 * NF + idN];
          }
      /* du_grad (ops.f90:787-832) */
      const double* m = mass[l] + (int64_t)k * NC;
      const double* t = temp[l] + (int64_t)k * NC;
      double* du = dvelo[l] + (int64_t)k * 3 * NC;
      const double* len = gc->len;
      for (int lz = 0; lz < 10; ++lz)
        for (int j = 0; j < gc->n; ++j)
          for (int i = 0; i < gc->n; ++i) {
            const int64_t id = cell_of(gc, lz, i, j), idE = cell_of(gc, lz, i + 1, j), idNE = cell_of(gc, lz, i + 1, j + 1),
                          idN = cell_of(gc, lz, i, j + 1);
            const double th0 = t[id] / m[id], thE = t[idE] / m[idE], thNE = t[idNE] / m[idNE], thN = t[idN] / m[idN];
            const double the[3] = {0.5 * (th0 + thE), 0.5 * (th0 + thNE), 0.5 * (th0 + thN)};
            const double* B = bern[l];
            const double* X = exner[l];
            const double gB[3] = {(B[idE] - B[id]) / len[id], (B[id] - B[idNE]) / len[NC + id],
                                  (B[idN] - B[id]) / len[2 * NC + id]};
            const double gE[3] = {(X[idE] - X[id]) / len[id], (X[id] - X[idNE]) / len[NC + id],
                                  (X[idN] - X[id]) / len[2 * NC + id]};
            for (int e = 0; e < 3; ++e)
              du[e * NC + id] = src[l][e * NC + id] / len[e * NC + id] - gB[e] - the[e] * gE[e];
          }
    }
  }
  for (int l = 0; l < top; ++l) {
    free(hm[l]);
    free(ht[l]);
    free(bern[l]);
    free(exner[l]);
    free(src[l]);
  }
  free(keep.hflux_m);
  free(keep.hflux_t);
  free(keep.bern);
  free(keep.exner);
  free(keep.du_src);
}

/* ---------------------------------------------------------------------------
 * WT_after_step and RK4
 * ------------------------------------------------------------------------- */
void wto_wt_after_step(const wto_ml* h, int L, double** mass, double** temp, double** velo) {
  const int nl = h->nlevels;
  for (int l = 0; l < nl; ++l) wto_ghost_fill(h->grid[l], L, mass[l], temp[l], velo[l]);
  /* wavelets of every level from the stage's state, before any level is
   * reconstructed (compute_scalar_wavelets, compute_velo_wavelets) */
  double *wcm[16], *wct[16], *wcu[16];
  for (int l = 1; l < nl; ++l) {
    const wto_grid* gf = h->grid[l];
    const int64_t NF = gf->ncells, NC = h->grid[l - 1]->ncells;
    wcm[l] = calloc((size_t)NF * L, sizeof(double));
    wct[l] = calloc((size_t)NF * L, sizeof(double));
    wcu[l] = calloc((size_t)3 * NF * L, sizeof(double));
    for (int k = 0; k < L; ++k) {
      wto_scalar_wavelets(gf, h->ovl[l], mass[l] + (int64_t)k * NF, wcm[l] + (int64_t)k * NF);
      wto_scalar_wavelets(gf, h->ovl[l], temp[l] + (int64_t)k * NF, wct[l] + (int64_t)k * NF);
      wto_velo_wavelets(h, l, velo[l] + (int64_t)k * 3 * NF, velo[l - 1] + (int64_t)k * 3 * NC,
                        wcu[l] + (int64_t)k * 3 * NF);
    }
  }
  /* inverse_wavelet_transform, coarse to fine */
  for (int l = 1; l < nl; ++l) {
    const wto_grid* gf = h->grid[l];
    const wto_grid* gc = h->grid[l - 1];
    const int64_t NF = gf->ncells, NC = gc->ncells;
    for (int k = 0; k < L; ++k) {
      wto_prolong(gf, gc, h->ovl[l], mass[l - 1] + (int64_t)k * NC, wcm[l] + (int64_t)k * NF, mass[l] + (int64_t)k * NF);
      wto_prolong(gf, gc, h->ovl[l], temp[l - 1] + (int64_t)k * NC, wct[l] + (int64_t)k * NF, temp[l] + (int64_t)k * NF);
      wto_velo_reconstruct(h, l, velo[l - 1] + (int64_t)k * 3 * NC, wcu[l] + (int64_t)k * 3 * NF,
                           velo[l] + (int64_t)k * 3 * NF);
    }
    free(wcm[l]);
    free(wct[l]);
    free(wcu[l]);
  }
}

void wto_rk4_ml(const wto_ml* h, int L, double dt, int nsteps, double** mass, double** temp, double** velo) {
  const int nl = h->nlevels;
  double *m1[16], *t1[16], *u1[16], *dm[16], *dtm[16], *du[16];
  for (int l = 0; l < nl; ++l) {
    const int64_t NC = h->grid[l]->ncells;
    m1[l] = malloc(sizeof(double) * NC * L);
    t1[l] = malloc(sizeof(double) * NC * L);
    u1[l] = malloc(sizeof(double) * 3 * NC * L);
    dm[l] = malloc(sizeof(double) * NC * L);
    dtm[l] = malloc(sizeof(double) * NC * L);
    du[l] = malloc(sizeof(double) * 3 * NC * L);
    wto_ghost_fill(h->grid[l], L, mass[l], temp[l], velo[l]);
  }
  const double coef[4] = {dt / 4.0, dt / 3.0, dt / 2.0, dt};
  for (int step = 0; step < nsteps; ++step) {
    for (int l = 0; l < nl; ++l) {
      const int64_t NC = h->grid[l]->ncells;
      memcpy(m1[l], mass[l], sizeof(double) * NC * L);
      memcpy(t1[l], temp[l], sizeof(double) * NC * L);
      memcpy(u1[l], velo[l], sizeof(double) * 3 * NC * L);
    }
    for (int st = 0; st < 4; ++st) {
      double** ym = st == 0 ? mass : m1;
      double** yt = st == 0 ? temp : t1;
      double** yu = st == 0 ? velo : u1;
      wto_trend_ml(h, L, ym, yt, yu, dm, dtm, du);
      double** om = st == 3 ? mass : m1;
      double** ot = st == 3 ? temp : t1;
      double** ou = st == 3 ? velo : u1;
      const double hh = coef[st];
      for (int l = 0; l < nl; ++l) {
        const wto_grid* g = h->grid[l];
        const int64_t NC = g->ncells;
        for (int k = 0; k < L; ++k)
          for (int64_t c = 0; c < NC; ++c) {
            if (!prognostic(g, c)) continue;
            const int64_t x = (int64_t)k * NC + c;
            om[l][x] = mass[l][x] + hh * dm[l][x];
            ot[l][x] = temp[l][x] + hh * dtm[l][x];
            for (int e = 0; e < 3; ++e) {
              const int64_t y = ((int64_t)k * 3 + e) * NC + c;
              ou[l][y] = velo[l][y] + hh * du[l][y];
            }
          }
      }
      wto_wt_after_step(h, L, om, ot, ou);
    }
  }
  for (int l = 0; l < nl; ++l) {
    free(m1[l]);
    free(t1[l]);
    free(u1[l]);
    free(dm[l]);
    free(dtm[l]);
    free(du[l]);
  }
}

/* ---------------------------------------------------------------------------
 * RK4_split on the hierarchy (time_integr.f90:360-382, 444-478)
 * ------------------------------------------------------------------------- */
/* flux_divergence (barotropic_2d.f90:326-366) on every level: the finest
 * level's vertically integrated flux (step1 itype 4) and its divergence,
 * coarser levels from the restricted flux (cpt_or_restr_flux with dscalar =
 * the finer divergence); divF[l] ghost filled */
static void flux_divergence_ml(const wto_ml* h, int L, double* const* mass, double* const* velo, double** divF) {
  const int top = h->nlevels - 1;
  const wto_grid* gt = h->grid[top];
  const int64_t NT = gt->ncells;
  double* hf = calloc(3 * NT, sizeof(double));
  wto_flux_divergence_parts(gt, L, mass[top], velo[top], divF[top], hf);
  edge_ghost(gt, hf);
  double* fine = hf;
  for (int l = top - 1; l >= 0; --l) {
    const wto_grid* gc = h->grid[l];
    double* hc = calloc(3 * gc->ncells, sizeof(double));
    wto_flux_restrict(h, l + 1, fine, divF[l + 1], hc);
    memset(divF[l], 0, sizeof(double) * gc->ncells);
    divergence(gc, hc, divF[l]);
    /* cal_div is +div: undo the scalar trend's sign */
    for (int64_t c = 0; c < gc->ncells; ++c) divF[l][c] = -divF[l][c];
    wto_node_ghost(gc, divF[l]);
    free(fine);
    fine = hc;
  }
  free(fine);
}

/* grad_eta (barotropic_2d.f90:457-474) on every level: eta copied, its
 * coarse values on [0, n]^2 replaced by the finer level's at the coincident
 * node (cpt_or_restr_eta), then g grad on the level's edges */
static void grad_eta_ml(const wto_ml* h, double* const* eta, double** hf) {
  const int top = h->nlevels - 1;
  double* e2[16];
  for (int l = top; l >= 0; --l) {
    const wto_grid* g = h->grid[l];
    e2[l] = malloc(sizeof(double) * g->ncells);
    memcpy(e2[l], eta[l], sizeof(double) * g->ncells);
    if (l < top) {
      const wto_grid* gf = h->grid[l + 1];
      for (int lz = 0; lz < 10; ++lz)
        for (int j = 0; j <= g->n; ++j)
          for (int i = 0; i <= g->n; ++i) e2[l][cell_of(g, lz, i, j)] = e2[l + 1][cell_of(gf, lz, 2 * i, 2 * j)];
    }
    wto_grad_eta(g, e2[l], hf[l]);
  }
  for (int l = 0; l <= top; ++l) free(e2[l]);
}

int wto_rk_split_ml(const wto_ml* h, wto_mg* mg, int L, int nstages, double** mass, double** temp, double** velo,
                    double** eta, int* fmg_iters) {
  const int nl = h->nlevels;
  const double dt = mg->p.dt, th1 = mg->p.theta1, th2 = mg->p.theta2;
  double *m1[16], *t1[16], *u1[16], *dm[16], *dtm[16], *du[16], *hf[16], *dF[16], *dFold[16], *rhs[16], *eo[16];
  for (int l = 0; l < nl; ++l) {
    const wto_grid* g = h->grid[l];
    const int64_t NC = g->ncells;
    m1[l] = malloc(8 * NC * L);
    t1[l] = malloc(8 * NC * L);
    u1[l] = malloc(8 * 3 * NC * L);
    dm[l] = malloc(8 * NC * L);
    dtm[l] = malloc(8 * NC * L);
    du[l] = malloc(8 * 3 * NC * L);
    hf[l] = calloc(3 * NC, 8);
    dF[l] = calloc(NC, 8);
    dFold[l] = calloc(NC, 8);
    rhs[l] = calloc(NC, 8);
    eo[l] = malloc(8 * NC);
    wto_ghost_fill(g, L, mass[l], temp[l], velo[l]);
    wto_node_ghost(g, eta[l]);
  }
  if (th2 != 1.0) flux_divergence_ml(h, L, mass, velo, dFold);
  for (int l = 0; l < nl; ++l) {
    const int64_t NC = h->grid[l]->ncells;
    memcpy(m1[l], mass[l], 8 * NC * L);
    memcpy(t1[l], temp[l], 8 * NC * L);
    memcpy(u1[l], velo[l], 8 * 3 * NC * L);
  }
  grad_eta_ml(h, eta, hf); /* eta is fixed through the stages */
  for (int st = 0; st < nstages; ++st) {
    double** ym = st == 0 ? mass : m1;
    double** yt = st == 0 ? temp : t1;
    double** yu = st == 0 ? velo : u1;
    for (int l = 0; l < nl; ++l) wto_barotropic_correction(h->grid[l], L, eta[l], ym[l], yt[l]);
    wto_trend_ml(h, L, ym, yt, yu, dm, dtm, du);
    const int last = st == nstages - 1;
    double** om = last ? mass : m1;
    double** ot = last ? temp : t1;
    double** ou = last ? velo : u1;
    const double hh = dt / (double)(nstages - st); /* h/4, h/3, h/2, h */
    for (int l = 0; l < nl; ++l) {
      const wto_grid* g = h->grid[l];
      const int64_t NC = g->ncells;
      for (int k = 0; k < L; ++k)
        for (int64_t c = 0; c < NC; ++c) {
          if (!prognostic(g, c)) continue;
          const int64_t x = (int64_t)k * NC + c;
          om[l][x] = mass[l][x] + hh * dm[l][x];
          ot[l][x] = temp[l][x] + hh * dtm[l][x];
          for (int e = 0; e < 3; ++e) {
            const int64_t y = ((int64_t)k * 3 + e) * NC + c;
            ou[l][y] = velo[l][y] + hh * (du[l][y] + th1 * hf[l][e * NC + c]);
          }
        }
    }
    wto_wt_after_step(h, L, om, ot, ou);
  }
  /* free_surface_update: eta_update (rhs_elliptic, FMG), barotropic
   * correction, u_update, WT_after_step with the free surface */
  flux_divergence_ml(h, L, mass, velo, dF);
  for (int l = 0; l < nl; ++l) {
    wto_rhs_elliptic(h->grid[l], &mg->p, h->grid[l]->rho0, eta[l], dF[l], dFold[l], rhs[l]);
    memcpy(eo[l], eta[l], 8 * h->grid[l]->ncells);
    mg->eta_old[l] = eo[l];
  }
  double res[16];
  wto_fmg(mg, eta, rhs, fmg_iters, res);
  for (int l = 0; l < nl; ++l) wto_barotropic_correction(h->grid[l], L, eta[l], mass[l], temp[l]);
  grad_eta_ml(h, eta, hf);
  for (int l = 0; l < nl; ++l) {
    const wto_grid* g = h->grid[l];
    const int64_t NC = g->ncells;
    for (int k = 0; k < L; ++k)
      for (int64_t c = 0; c < NC; ++c) {
        if (!prognostic(g, c)) continue;
        for (int e = 0; e < 3; ++e) {
          const int64_t y = ((int64_t)k * 3 + e) * NC + c;
          velo[l][y] = velo[l][y] - th1 * dt * hf[l][e * NC + c];
        }
      }
  }
  /* the free surface's own transform (the extra layer of WT_after_step) */
  double* wce[16];
  for (int l = 1; l < nl; ++l) {
    wce[l] = calloc(h->grid[l]->ncells, 8);
    wto_node_ghost(h->grid[l], eta[l]);
    wto_scalar_wavelets(h->grid[l], h->ovl[l], eta[l], wce[l]);
  }
  wto_wt_after_step(h, L, mass, temp, velo);
  for (int l = 1; l < nl; ++l) {
    wto_prolong(h->grid[l], h->grid[l - 1], h->ovl[l], eta[l - 1], wce[l], eta[l]);
    free(wce[l]);
  }
  for (int l = 0; l < nl; ++l) {
    free(m1[l]); free(t1[l]); free(u1[l]); free(dm[l]); free(dtm[l]); free(du[l]);
    free(hf[l]); free(dF[l]); free(dFold[l]); free(rhs[l]); free(eo[l]);
  }
  return 0;
}

int wto_rk4_split_ml(const wto_ml* h, wto_mg* mg, int L, double** mass, double** temp, double** velo,
                     double** eta, int* fmg_iters) {
  return wto_rk_split_ml(h, mg, L, 4, mass, temp, velo, eta, fmg_iters);
}
