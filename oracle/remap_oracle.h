/* CPU oracle of the vertical remap of Lagrangian layers (test infrastructure
 * only: tests, smoke and the bench's CPU baseline use it as the checker). */
#ifndef WTO_REMAP_ORACLE_H
#define WTO_REMAP_ORACLE_H
#include "trisk_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* one column: var_new on the cells of z_new from var_old on the cells of
 * z_old (both [0..N] interface heights, bottom up); scheme 0 = remap0,
 * 1 = remap1, 2 = remap2PPM (remap.f90:259-495), 3 = remapPPR
 * (remap.f90:950-1010, order 2 with the mono limiter and loose ends),
 * 4 = remap2S, 5 = remap2W, 6 = remap4 (remap.f90:497-948) */
void wto_remap_column(int scheme, int N, double* var_new, const double* z_new, const double* var_old,
                      const double* z_old);

/* remap_vertical_coordinates, incompressible branch (remap.f90:18-75,
 * 147-257) with sol_mean = 0, no porosity and the hybrid coordinates
 * z_new(k) = a_vert(k) eta + b_vert(k) z_s (a_vert, b_vert [L + 1]): every
 * owned column and the two pole masters (their velocities are left as they
 * are), then the ghost fill.  The state must be ghost filled on entry. */
void wto_remap(const wto_grid* g, int scheme, int L, const double* a_vert, const double* b_vert, double* mass,
               double* temp, double* velo);

#ifdef __cplusplus
}
#endif
#endif
