/* CPU oracle of the multi-level scalar transforms and FMG (see mg_oracle.c).
 * TEST INFRASTRUCTURE ONLY. */
#ifndef WTG_MG_ORACLE_H
#define WTG_MG_ORACLE_H
#include "elliptic_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* levels 0 .. nlevels-1, coarse to fine; ovl[l] = overlap areas between
 * levels l and l-1 ([4][ncells of level l], ovl[0] unused) */
typedef struct {
  int nlevels;
  const wto_grid* grid[32];
  const double* ovl[32];
  const double* eta_old[32];
  const double* weight[32];
  wto_fs_params p;
  double coarse_tol, fine_tol;
  int coarse_iter, max_vcycle, pre_iter, down_iter, up_iter, post_iter;
} wto_mg;

/* wc at the fine midpoint cells from the fine scalar s */
void wto_scalar_wavelets(const wto_grid* gf, const double* ovl, const double* s, double* wc);
/* coarse scalar from the fine scalar and its wavelets (sub-sample and lift) */
void wto_restrict(const wto_grid* gf, const wto_grid* gc, const double* ovl, const double* s_f, const double* wc,
                  double* s_c);
/* fine scalar from the coarse scalar and the fine wavelets */
void wto_prolong(const wto_grid* gf, const wto_grid* gc, const double* ovl, const double* s_c, const double* wc,
                 double* s_f);
/* full multigrid solve of elliptic_lo u = f; u[l], f[l] per level, ghost
 * filled; iters[l] = BiCGSTAB iterations (l = 0) or V-cycles, res[l] the
 * final relative residual per level */
void wto_fmg(const wto_mg* h, double** u, double* const* f, int* iters, double* res);
/* SRJ multiscale solve (same arguments; iters[l>0] = Jacobi sweeps) */
void wto_srj(const wto_mg* h, int max_srj_iter, double** u, double* const* f, int* iters, double* res);

#ifdef __cplusplus
}
#endif
#endif
