/* wavetrisk_gpu.h -- C ABI of the MI355X dynamical-core hot path.
 *
 * Drop-in boundary for the procedures WAVETRISK's time integrator calls:
 *   trend_fun => trend_ml        (time_integr.f90:7-13, multi_level.f90:5-47)
 *   dt_step   => RK4             (time_integr.f90:15-31, 78-103; main.f90:411)
 *   update_bdry                  (comm_ser.f90:113-152 -> comm.f90:785-808)
 * for a uniform level of the icosahedral grid.  A Fortran caller binds these
 * with ISO_C_BINDING; all arguments are plain pointers,
 * sizes and scalars.  Every function returns 0 on success and a negative
 * WTG_E_* code on failure (the Fortran shim calls abort on non-zero, as the
 * reference's `call abort` does); wtg_last_error() gives the message.
 *
 * Data layout (host arrays passed in and out).  The sphere is 10 lozenges
 * of n x n nodes, n = 2^level, each stored as a tile of S x S cells,
 * S = n + 2*halo, halo = 2; cell (lz, i, j) = lz*S*S + (j+halo)*S + (i+halo)
 * for i, j in [-halo, n+halo).  Owned nodes have i, j in [0, n); the south
 * pole is cell (0, n, 0), the north pole cell (5, 0, n).  A node owns edges
 * RT (i,j)->(i+1,j), DG (i+1,j+1)->(i,j), UP (i,j)->(i,j+1).
 *   scalar fields (mass = rho*dz, temp = rho*dz*theta): [nlayers][ncells]
 *   velocity (edge-normal, m/s):                        [nlayers][3][ncells]
 * Layer 0 is the bottom layer (k = 1 in the reference).
 */
#ifndef WAVETRISK_GPU_H
#define WAVETRISK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WTG_OK = 0,
  WTG_E_ARG = -1,      /* bad argument */
  WTG_E_DEVICE = -2,   /* HIP runtime error (no GPU, out of memory, launch failure) */
  WTG_E_STATE = -3,    /* call out of order (e.g. topography not set) */
  WTG_E_INTERNAL = -4
};

/* Grid: the reference's compile-time parameters (src/param_J*.f90) plus
 * the level.  min_level and domain_level fix how node coordinates round and
 * one ghost-list convention beside the south pole, so they must match the
 * reference build being replaced. */
typedef struct {
  int level;         /* J: n = 2^J nodes per lozenge side */
  int min_level;     /* MIN_LEVEL of the reference build, min_level-1 <= level */
  int domain_level;  /* DOMAIN_LEVEL of the reference build */
  double radius;     /* m */
  double omega;      /* s^-1 */
} wtg_grid_desc;

typedef struct {
  int n, S, halo;
  int64_t ncells;      /* 10*S*S */
  int64_t south_pole;  /* cell index of the south pole */
  int64_t north_pole;  /* cell index of the north pole */
} wtg_grid_info;

/* Dynamics parameters of the Boussinesq (incompressible) layer model. */
typedef struct {
  int nlayers;         /* zlevels */
  double grav;         /* grav_accel, m s^-2 */
  double ref_density;  /* ref_density, kg m^-3 */
  double p_top;        /* pressure at the top of the top layer */
} wtg_params;

/* ---- host-side grid (no GPU needed) ------------------------------------ */
typedef struct wtg_grid wtg_grid;

/* Builds coordinates, ghost maps and TRiSK geometry of one level; replaces
 * init_grid / precompute_geometry / add_second_level (init.f90:142-524,
 * refine_patch.f90:102-186) for a uniform level. */
int wtg_grid_create(const wtg_grid_desc* desc, wtg_grid** out);
int wtg_grid_info_get(const wtg_grid* grid, wtg_grid_info* info);
/* Copies one host array of the grid.  name: "x" [3][ncells] coordinates,
 * "len", "pedlen" [3][ncells], "area" [6][ncells], "hex_inv" [ncells],
 * "triarea", "coriolis" [2][ncells] (double); "node_src" [ncells],
 * "edge_src" [3*ncells] (int64; -1 = frozen halo slot, see below);
 * "edge_sign" [3*ncells] (int8); "wide_src" [10*(n+10)^2] (int64): the
 * owned cell of every lattice point (lz, i, j), i, j in [-5, n+5), index
 * lz*(n+10)^2 + (j+5)*(n+10) + i+5, -1 where the lattice does not continue.
 * bytes must equal the array's size. */
int wtg_grid_get(const wtg_grid* grid, const char* name, void* out, size_t bytes);
/* Overlap areas of the scalar wavelet transform between a grid and the next
 * coarser one (get_overl_areas / set_RF_wgts, wavelet.f90:1408-1539):
 * [4][fine ncells] doubles, a(1..4) at fine nodes that are coarse edge
 * midpoints, 0 elsewhere.  Both grids need the same min_level/domain_level. */
int wtg_grid_overlap_areas(const wtg_grid* fine, const wtg_grid* coarse, double* out, size_t bytes);
/* Inter-level weights of the velocity and flux transforms between a grid and
 * the next coarser one, as the reference computes them when it refines
 * (set_WT_wgts / outer_velo_weights, set_RF_wgts / basic_F_restr_wgt,
 * wavelet.f90:1176-1360, 1512-1705; replaces their host computation):
 *   "split" [2][fine ncells]  split parts of the overlap areas
 *   "iu"    [9][fine ncells]  I_u_wgt at the second end point of each outer child edge
 *   "rf"    [4][6][coarse ncells]  R_F_wgt per coarse node: RT side k=0, RT k=1,
 *           UP k=0, UP k=1, two triplets each (id_enc order). */
int wtg_grid_transfer_weights(const wtg_grid* fine, const wtg_grid* coarse, const char* name, double* out,
                              size_t bytes);
void wtg_grid_destroy(wtg_grid* grid);

/* ---- device context ------------------------------------------------------ */
typedef struct wtg_ctx wtg_ctx;

/* Uploads the grid to the current HIP device and allocates nslots user
 * state slots (>= 1) plus two internal stage buffers.  One HIP stream per
 * context; calls are stream ordered and the host functions below
 * synchronise before returning. */
int wtg_create(const wtg_grid* grid, const wtg_params* params, int nslots, wtg_ctx** out);
/* An ensemble of nmembers independent states on the same grid (the batched
 * ensemble of config 5): every slot holds the members one after another,
 * scalars [member][layer][cell], velocity [member][layer][edge][cell], and
 * every layer kernel runs all members in one launch (members on the grid's
 * second dimension); each member has its own topography.  wtg_trend,
 * wtg_rk4, wtg_rk4_split and the ghost fill take the whole ensemble (the
 * free-surface fields then hold [member][cell]); the other free-surface
 * calls and hierarchies need nmembers = 1. */
int wtg_create_ensemble(const wtg_grid* grid, const wtg_params* params, int nmembers, int nslots, wtg_ctx** out);
void wtg_destroy(wtg_ctx* ctx);

/* Bottom height z_s per cell [ncells] (topography, m), [nmembers][ncells]
 * for an ensemble; halo cells may be anything, they are filled from their
 * owners. */
int wtg_set_topography(wtg_ctx* ctx, const double* topo);

/* Host <-> device copies of a state slot (layouts above).  Halo slots whose
 * ghost map is -1 ("frozen": the reference's ghost lists never refresh them
 * and a time step never changes them) keep the uploaded values; every other
 * halo slot is refilled from its owner on upload. */
int wtg_upload_state(wtg_ctx* ctx, int slot, const double* mass, const double* temp, const double* velo);
int wtg_download_state(wtg_ctx* ctx, int slot, double* mass, double* temp, double* velo);

/* update_bdry on one slot: halo := owner values (edges with sign). */
int wtg_ghost_fill(wtg_ctx* ctx, int slot);

/* Sharding model (no counterpart in the reference; one GPU runs it): shards
 * of whole lozenges, lozenge lz in shard lz * nshards / 10 (the reference's
 * 16 domains of a lozenge together, comm.f90's lists cut at lozenge edges).
 * wtg_grid_shard_halo counts the ghost entries of one layer: counts[0] node
 * entries whose source is in the destination's shard, counts[1] the others
 * (the would-be remote ones), counts[2], counts[3] the same for edges.
 * wtg_ghost_fill_split_timed fills slot `slot` as those two launches, each
 * timed over `reps` launches (mean ms); the result equals wtg_ghost_fill. */
int wtg_grid_shard_halo(const wtg_grid* grid, int nshards, int64_t counts[4]);
int wtg_ghost_fill_split_timed(wtg_ctx* ctx, int slot, int nshards, int reps, double* ms_local, double* ms_remote);

/* trend_ml: dq_slot := trend of q_slot (q_slot must be ghost filled).
 * The trend is written on owned cells/edges and the pole masters. */
int wtg_trend(wtg_ctx* ctx, int q_slot, int dq_slot);

/* dt_step => RK4: nsteps low-storage RK4 steps of size dt on slot `slot`
 * (q1 = q + dt/4 f(q); q1 = q + dt/3 f(q1); q1 = q + dt/2 f(q1);
 * q = q + dt f(q1), each stage followed by a ghost fill). */
int wtg_rk4(wtg_ctx* ctx, int slot, double dt, int nsteps);

/* Same as wtg_rk4, timed with HIP events on the context's stream:
 * total_ms for all steps, stage_kernel_ms = mean duration of one launch of
 * the fused trend+update kernel (the dominant kernel). */
int wtg_rk4_timed(wtg_ctx* ctx, int slot, double dt, int nsteps, double* total_ms, double* stage_kernel_ms);

/* Precision of wtg_rk4 and wtg_rk4_timed on this context (no counterpart in
 * the reference, which is FP64 throughout):
 *   WTG_PRECISION_FP64      the FP64 path, bit-exact with trend_ml / RK4
 *                           (time_integr.f90:78-103, multi_level.f90:5-47);
 *   WTG_PRECISION_FP32_BULK the same step with potential vorticity, fluxes,
 *                           kinetic energy, divergence and Qperp in FP32 and
 *                           the intermediate RK stages kept as FP32
 *                           increments over the FP64 state; the state, the
 *                           hydrostatic column scan, Bernoulli / Exner
 *                           gradients and the update of the state stay FP64.
 *                           Its error against the FP64 path is bounded by the
 *                           tests, not bit-exact.
 * The first switch to FP32_BULK allocates two FP32 increment buffers and FP32
 * geometry copies.  Other entry points (wtg_trend, the split step) stay FP64. */
#define WTG_PRECISION_FP64 0
#define WTG_PRECISION_FP32_BULK 1
int wtg_set_precision(wtg_ctx* ctx, int precision);

/* Launch shape of the FP32-bulk stage kernel (no counterpart in the
 * reference): WTG_SHAPE_LAYER_PAIRS (one 768-thread block per 16 x 16 tile,
 * two layers in lock step, one owned cell per thread in the trend phase),
 * WTG_SHAPE_EDGE_TASKS (the same with the trend phase split into edge tasks
 * over all threads), WTG_SHAPE_ONE_LAYER (384 threads, one layer at a time,
 * three blocks per CU), or WTG_SHAPE_AUTO (default): the first FP32-bulk step
 * times every shape on the context's state and keeps the fastest.  All shapes
 * compute the same arithmetic per cell.  wtg_get_stage_shape reports the
 * shape in use (WTG_SHAPE_AUTO before tuning) and, in ms[0..2], the tuner's
 * mean launch time of shapes 1..3 (0 when the shape was set by hand). */
#define WTG_SHAPE_AUTO 0
#define WTG_SHAPE_LAYER_PAIRS 1
#define WTG_SHAPE_EDGE_TASKS 2
#define WTG_SHAPE_ONE_LAYER 3
int wtg_set_stage_shape(wtg_ctx* ctx, int shape);
int wtg_get_stage_shape(wtg_ctx* ctx, int* shape, float* ms);

/* enable != 0: wtg_rk4 replays its steps from a HIP graph captured on the
 * context's stream (one step, or two in FP32-bulk mode, per graph; captured
 * on first use and again when slot, dt, precision or launch shape change).
 * Results are identical to the eager launches.  No counterpart in the
 * reference.  wtg_rk4_timed always launches eagerly (it times each stage). */
int wtg_set_graph(wtg_ctx* ctx, int enable);

/* Laplacian diffusion (order 1) or iterated hyperdiffusion (order 2) in the
 * trend of wtg_trend, wtg_rk4 and the split-explicit steps; replaces the
 * reference's Laplace_sclr / Laplace_divu / Laplace_rotu, visc_sclr(S_MASS),
 * visc_sclr(S_TEMP), visc_divu, visc_rotu (shared.f90:226, 238) with its test
 * cases' physics hooks (physics_scalar_flux, physics_velo_source;
 * ops.f90:360, 511, 778; e.g. test/seamount/test_case_module.f90:106-156):
 *   mass / temperature flux += (-1)^Ls visc grad(q or Laplacian q) pedlen
 *   velocity source          += (-1)^(Lr-1) (visc_divu grad(div u or its
 *                               Laplacian) - visc_rotu curl(vorticity or
 *                               rot rot of it))
 * computed as trend_ml does (cal_divu_ml, cal_Laplacian_scalars,
 * cal_Laplacian_divu, cal_Laplacian_vector_rot; multi_level.f90:26-30, 99,
 * 190-309, 579-617).  Orders are 0 (off), 1 or 2; the velocity term is on
 * when laplace_rotu is, and laplace_divu needs laplace_rotu (the hook's
 * sign).  d == NULL or all orders 0 switches diffusion off.  FP64 only: the
 * FP32-bulk precision, the stage kernels other than the default, and the
 * multi-level step refuse a context with diffusion on.  The first call
 * allocates the pre-pass buffers (up to 14 fields of [layers][ncells]). */
typedef struct {
  int laplace_sclr, laplace_divu, laplace_rotu;
  double visc_mass, visc_temp, visc_divu, visc_rotu;
} wtg_diffusion;
int wtg_set_diffusion(wtg_ctx* ctx, const wtg_diffusion* d);

/* remap_vertical_coordinates (remap.f90:18-75) of the incompressible layer
 * model (remap_incompressible, remap.f90:147-216): the Lagrangian layers of
 * slot `slot` remapped onto the hybrid coordinates z(k) = a_vert[k] eta +
 * b_vert[k] z_s (k = 0..nlayers, bottom up, host arrays; the reference's
 * a_vert / b_vert after initialize_a_b_vert), every member of an ensemble:
 * the buoyancy and the edge velocities (not at the poles) remapped with the
 * reference's remap_type "0" (WTG_REMAP_PIECEWISE_CONSTANT, remap0), "1"
 * (WTG_REMAP_PIECEWISE_LINEAR, remap1), "2PPM" (WTG_REMAP_PARABOLIC_PPM,
 * remap2PPM), its default "PPR" (WTG_REMAP_PPR, remapPPR at
 * remap.f90:950-1010: order 2, mono limiter, loose ends), "2S"
 * (WTG_REMAP_PARABOLIC_SPLINE, remap2S; at least 2 layers), "2W"
 * (WTG_REMAP_PARABOLIC_WENO, remap2W) or "4" (WTG_REMAP_QUARTIC_WENO,
 * remap4; at least 3 layers), the new mass rho0 dz; the slot is ghost
 * filled before and after.  The reference's sigma_z coordinates and its
 * compressible branch are not built.  At most WTG_REMAP_MAX_LAYERS layers;
 * the first work slot holds the old masses meanwhile. */
#define WTG_REMAP_PIECEWISE_CONSTANT 0
#define WTG_REMAP_PIECEWISE_LINEAR 1
#define WTG_REMAP_PARABOLIC_PPM 2
#define WTG_REMAP_PPR 3
#define WTG_REMAP_PARABOLIC_SPLINE 4
#define WTG_REMAP_PARABOLIC_WENO 5
#define WTG_REMAP_QUARTIC_WENO 6
#define WTG_REMAP_MAX_LAYERS 64
int wtg_remap(wtg_ctx* ctx, int slot, int scheme, const double* a_vert, const double* b_vert);

/* vertical_diffusion (vert_diffusion.f90:66-86), the ocean's implicit
 * vertical mixing after the dynamics of a step (main.f90:420): per column the
 * eddy diffusivity Kt and viscosity Kv at the layer interfaces, from the
 * analytic profiles (Kt_analytic, Kv_analytic) or the one-equation TKE closure
 * (turbulent_diffusion: a backward-Euler TKE step through a tridiagonal
 * solve, then update_Kv_Kt), then backward-Euler steps of the buoyancy
 * (backwards_euler_temp) and of the three edge velocities
 * (backwards_euler_velo), each a tridiagonal solve (LAPACK dgtsv, with its
 * partial pivoting).  The fields of wtg_vdiff are vert_diffusion_mod's
 * module variables (vert_diffusion.f90:27-64), shared.f90's tke_closure and
 * c_p, and the test case's bottom_friction; wtg_vdiff_defaults fills the
 * reference's defaults (bottom_friction 0).  mode_split 1 takes eta from the
 * free-surface field (sol(S_MASS, zlevels+1) of the split-explicit step), 0
 * from the column (free_surface, utils.f90:505-526). */
typedef struct {
  int tke_closure, patankar, enhance_diff, mode_split;
  double C_e, C_eps, C_l, C_k, C_srf, e_0, e_min, e_min_srf, eps_s, kappa_VK, Kt_enh, Kt_max, Kt_min, Kv_max,
      Kv_min, mixed_layer, l_0, l_min, Neps_sq, Nsq_min, Q_sr, c_p, bottom_friction;
} wtg_vdiff;
void wtg_vdiff_defaults(wtg_vdiff* p);
/* enables the vertical diffusion with p (at least 2 layers, 3 with the TKE
 * closure); allocates Kt, Kv and the TKE (zero, as init.f90:513-520 leaves
 * them) and the forcing (zero) on first use */
int wtg_set_vdiff(wtg_ctx* ctx, const wtg_vdiff* p);
/* host arrays, every member one after another: Kt, Kv [nlayers + 1][ncells]
 * (interfaces 0..nlayers, bottom up), tke [nlayers][ncells] (layers 1..nlayers) */
int wtg_vdiff_set_coefficients(wtg_ctx* ctx, const double* Kt, const double* Kv, const double* tke);
int wtg_vdiff_get_coefficients(wtg_ctx* ctx, double* Kt, double* Kv, double* tke);
/* the test case's hooks evaluated per node (host arrays, members one after
 * another): bottom_buoy_flux, top_buoy_flux, tau_mag [ncells] and wind_flux
 * [3][ncells] (the RT, DG, UP edges of each node); NULL keeps an array */
int wtg_vdiff_set_forcing(wtg_ctx* ctx, const double* bottom_buoy_flux, const double* top_buoy_flux,
                          const double* tau_mag, const double* wind_flux);
/* one call of vertical_diffusion on slot with time step dt; istep 1 (the
 * first step of a run, main.f90:384) recomputes Kt and Kv from the TKE
 * before its step, as init_diffuse does.  The slot is ghost filled before and
 * after. */
int wtg_vertical_diffusion(wtg_ctx* ctx, int slot, double dt, int istep);

/* Device addresses of a slot's arrays (for zero-copy interop). */
int wtg_state_device_ptrs(wtg_ctx* ctx, int slot, double** mass, double** temp, double** velo);

int wtg_synchronize(wtg_ctx* ctx);

/* Last error message of ctx (or of the last failed create when ctx is NULL). */
int wtg_last_error(const wtg_ctx* ctx, char* buf, size_t n);

/* ---------------------------------------------------------------------------
 * Free-surface (barotropic) elliptic problem of the mode-split step, one
 * uniform level (barotropic_2d.f90, lin_solve.f90).  Node fields of the
 * context, each [ncells] like a one-layer scalar:
 *   WTG_FS_ETA      free-surface perturbation   (sol(S_MASS, zlevels+1))
 *   WTG_FS_RHS      right-hand side             (sol(S_TEMP, zlevels+1))
 *   WTG_FS_ETA_OLD  free surface in the operator (Laplacian_scalar(S_TEMP))
 *   WTG_FS_DIVF     flux divergence of the layer-integrated velocity
 *                   (trend(S_MASS, zlevels+1))
 *   WTG_FS_DIVF_OLD the same at the previous step (trend(S_TEMP, zlevels+1))
 * Porosity is 1 and every node is active.  Fields are ghost filled by every
 * call that writes them.
 * ------------------------------------------------------------------------- */
enum { WTG_FS_ETA = 0, WTG_FS_RHS = 1, WTG_FS_ETA_OLD = 2, WTG_FS_DIVF = 3, WTG_FS_DIVF_OLD = 4 };
enum { WTG_SOLVE_BICGSTAB = 0, WTG_SOLVE_JACOBI = 1 };

typedef struct {
  double dt;     /* time step (dt of shared.f90) */
  double theta1; /* implicitness of the external pressure gradient (theta1, shared.f90:437) */
  double theta2; /* implicitness of the barotropic divergence (theta2, shared.f90:438) */
} wtg_fs_params;

int wtg_fs_set_params(wtg_ctx* ctx, const wtg_fs_params* p);
/* host <-> device copy of one free-surface field ([ncells] doubles, or
 * [nmembers][ncells] for an ensemble; halo cells are filled from their
 * owners on upload) */
int wtg_fs_upload(wtg_ctx* ctx, int field, const double* host);
int wtg_fs_download(wtg_ctx* ctx, int field, double* host);

/* out = elliptic_lo(in) (barotropic_2d.f90:187-253):
 *   theta1 theta2 dt^2 div(g (|topo| + eta_old) grad in) - in */
int wtg_elliptic_apply(wtg_ctx* ctx, int in_field, int out_field);
/* out = elliptic_lo_diag (barotropic_2d.f90:255-324, the averaged-depth branch) */
int wtg_elliptic_diag(wtg_ctx* ctx, int out_field);
/* nsweeps of Jacobi_iteration (lin_solve.f90:271-340): u += w (f - L u) / diag */
int wtg_jacobi(wtg_ctx* ctx, int u_field, int f_field, double w, int nsweeps);
/* dp (lin_solve.f90:477-503): sum over the reference's (0,1) node sweep;
 * l2 (lin_solve.f90:433-461) is sqrt(dot(a, a)) */
int wtg_fs_dot(wtg_ctx* ctx, int a_field, int b_field, double* out);
/* wtg_fs_dot for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops); no counterpart in the reference.  nd (1..3)
 * weighted dot products sum_c w_c a[k][c] b[k][c] over the level's
 * reduction cells, in wtg_fs_dot's order (bit-identical results): a[k], b[k]
 * are device arrays [ncells] of the context's level, out a device array of nd
 * doubles.  Queued on the context's stream, no synchronisation; allocates
 * nothing (no hipMalloc, no free-surface fields).  The block partials live in
 * one file-scope device array per GPU (12 KiB, 512 partials of 3 sums, kept
 * until the process ends): calls that overlap on one GPU race.  nd > 3 is
 * refused. */
int wtg_fs_dots_noalloc(wtg_ctx* ctx, int nd, const double* const* a, const double* const* b, double* out);
/* WTG_FS_DIVF := flux_divergence of state slot `slot` (barotropic_2d.f90:326-366) */
int wtg_flux_divergence(wtg_ctx* ctx, int slot);
/* WTG_FS_RHS := -eta + dt (theta2 divF + (1 - theta2) divF_old) / ref_density
 * (rhs_elliptic, barotropic_2d.f90:151-185) */
int wtg_rhs_elliptic(wtg_ctx* ctx);
/* Solves elliptic_lo(eta) = rhs for WTG_FS_ETA in place, from its current
 * value, until ||rhs - L eta|| / ||rhs|| <= rel_tol (norms as l2 above) or
 * max_iter iterations.  WTG_SOLVE_BICGSTAB is bicgstab (lin_solve.f90:361-431),
 * the reference's solver on the coarsest level (and on the only level of a
 * uniform grid, FMG / SRJ lin_solve.f90:33-231); WTG_SOLVE_JACOBI is
 * Jacobi_iteration with weight 1.  *iters as the reference's iter_out
 * (max_iter + 1 if rel_tol was not reached), *rel_resid the last relative
 * residual. */
int wtg_elliptic_solve(wtg_ctx* ctx, int method, double rel_tol, int max_iter, int* iters, double* rel_resid);
/* enable != 0: BiCGSTAB replays each full chunk of its gated iterations
 * (8 iterations, 96 kernels) from a HIP graph captured on first use, instead
 * of launching the kernels one by one; the iterates and the iteration count
 * are those of the eager launches.  0 (default): eager launches.  No
 * counterpart in the reference. */
int wtg_fs_set_graph(wtg_ctx* ctx, int enable);

/* dt_step_split => RK4_split (time_integr.f90:360-382, main.f90:209): nsteps
 * split-explicit steps of size dt (wtg_fs_set_params) on state slot `slot`
 * with the free surface in WTG_FS_ETA (sol(S_MASS, zlevels+1)).  Per step:
 * flux divergence of the old state into WTG_FS_DIVF_OLD; four RK_split stages
 * (444-462: barotropic_correction of the stage input, trend_ml, scalar_star
 * and u_star with the theta1-weighted external pressure gradient removed);
 * free_surface_update (464-478: rhs_elliptic, eta_old := eta, the elliptic
 * solve of WTG_SOLVE_BICGSTAB to rel_tol -- the reference's SRJ on a single
 * level --, barotropic_correction, u_update).  Porosity 1, mean state 0.
 * *iters = BiCGSTAB iterations of the last step.  Ensembles (no counterpart
 * in the reference): one trend launch for all members, each member's
 * column and free-surface steps on its own fields, the members' BiCGSTAB
 * solves batched (one launch per step of the iteration for all members,
 * each member stopping at its own convergence test); *iters = the most any
 * member took in the last step. */
int wtg_rk4_split(wtg_ctx* ctx, int slot, int nsteps, double rel_tol, int max_iter, int* iters);
/* The same with nstages stages of RK_split: 4 = RK4_split, 3 = RK3_split
 * (time_integr.f90:384-405), 2 = RK2_split (:407-427), 1 = Euler_split
 * (:429-441); stage coefficients dt/nstages, ..., dt/2, dt as the
 * reference's low-storage schemes. */
int wtg_rk_split(wtg_ctx* ctx, int slot, int nstages, int nsteps, double rel_tol, int max_iter, int* iters);

/* ---- multi-level free surface: scalar wavelet transforms, FMG, SRJ -------
 * A hierarchy is a list of free-surface contexts on consecutive uniform
 * levels, coarse (index 0) to fine, each with its topography, dt / theta and
 * eta_old (WTG_FS_ETA_OLD); the reference solves on the levels
 * level_start..level_end of its grid with the same dt and theta on each.
 * While a hierarchy exists its contexts run on the finest context's stream;
 * destroy it before its contexts.  Errors of the hierarchy calls are
 * reported through wtg_last_error(finest context) (NULL for create).
 * ------------------------------------------------------------------------- */
typedef struct wtg_mg wtg_mg;
enum { WTG_SOLVE_FMG = 2, WTG_SOLVE_SRJ = 3, WTG_SOLVE_VCYCLE = 4, WTG_SOLVE_LEARNED = 5 };

/* lin_solve.f90:11-26 */
typedef struct {
  double coarse_tol;  /* BiCGSTAB tolerance on the coarsest level */
  double fine_tol;    /* relative residual that ends the work on a finer level */
  int coarse_iter;    /* BiCGSTAB iteration limit */
  int max_vcycle;     /* FMG: V-cycles per level */
  int pre_iter, down_iter, up_iter, post_iter;  /* FMG: Jacobi sweeps */
  int max_srj_iter;   /* SRJ: sweeps per level (rounded down to a multiple of 8) */
  int learned_cycles; /* WTG_SOLVE_LEARNED: the network's correction precedes the
                         first learned_cycles cycles only (no counterpart in the
                         reference; default: every cycle) */
} wtg_mg_params;

/* the reference's module defaults (coarse_tol = fine_tol = 1e-3, ...) */
void wtg_mg_default_params(wtg_mg_params* p);
/* grids[l] the grid levels[l] was created from; overlap areas between
 * consecutive levels are computed here (wavelet.f90:1408-1539). */
int wtg_mg_create(const wtg_grid* const* grids, wtg_ctx* const* levels, int nlevels, wtg_mg** out);
void wtg_mg_destroy(wtg_mg* mg);
/* elliptic_solver (init.f90:127) => FMG (lin_solve.f90:33-140) or SRJ
 * (lin_solve.f90:142-231): solves elliptic_lo(u) = f on every level, u in
 * each level's WTG_FS_ETA (initial guess in, solution out; a finer level's
 * guess is replaced by the prolongation of the coarser solution, as in the
 * reference), f in each level's WTG_FS_RHS.  iters[l]: BiCGSTAB iterations
 * on level 0, V-cycles (FMG) or Jacobi sweeps (SRJ) above; rel_resid[l] the
 * final ||f - L u|| / ||f|| per level ([nlevels] each). */
int wtg_mg_solve(wtg_mg* mg, int method, const wtg_mg_params* p, int* iters, double* rel_resid);
/* A hierarchy of ensemble contexts (wtg_create_ensemble, the same member
 * count on every level; no counterpart in the reference, which solves one
 * right-hand side at a time): wtg_mg_solve with WTG_SOLVE_VCYCLE or
 * WTG_SOLVE_LEARNED solves every member's problem ([nmembers][ncells] free-
 * surface fields) in the same launches -- one launch per operator of the
 * V-cycle, the members' coarse BiCGSTAB batched, one network inference for
 * all members' patches (the network's batch must hold npatches x nmembers
 * rows).  Each member's residual is tested after every cycle; a converged
 * member's fields are left alone from then on, so its iterates and cycle
 * count are those of its own solve.  iters / rel_resid of the finest level:
 * the most cycles and the largest residual; per member: */
int wtg_mg_member_results(const wtg_mg* mg, int* cycles, double* rel_resid);
/* WTG_SOLVE_VCYCLE: plain V-cycles (v_cycle, lin_solve.f90:102-139) on the
 * finest level from its current u until fine_tol or max_vcycle cycles.
 * WTG_SOLVE_LEARNED (no counterpart in the reference): the same cycles, each
 * preceded by the attached network's correction e of the residual r, taken
 * with the step a = <r, L e> / <L e, L e> when a > 0; the V-cycle is the
 * fallback and the FP64 residual ends the iteration.  For both, iters and
 * rel_resid report the finest level (coarser entries 0). */
/* wtg_rk4_split on the finest level's context with the free surface solved
 * on the hierarchy: eta^n restricted to every coarser level as its
 * operator's eta_old, then WTG_SOLVE_VCYCLE or WTG_SOLVE_LEARNED cycles from
 * eta^n to p->fine_tol (at most p->max_vcycle).  *iters = cycles of the
 * last step. */
int wtg_mg_rk4_split(wtg_mg* mg, int slot, int nsteps, int method, const wtg_mg_params* p, int* iters);
/* Precision of the V-cycles of WTG_SOLVE_FMG, WTG_SOLVE_VCYCLE and
 * WTG_SOLVE_LEARNED (the reference's v_cycle, lin_solve.f90:102-139, is
 * FP64 only).  WTG_PRECISION_FP64 (the default) runs them as the reference;
 * WTG_PRECISION_FP32_BULK makes each cycle a defect correction: the FP64
 * residual R = f - L u, its correction equation L e = R smoothed, restricted
 * and prolonged in FP32 fields, the coarsest level's BiCGSTAB in FP64, and
 * u := u + e and the residual that judges convergence in FP64.  The FP32
 * fields are allocated on every level at the first call with this mode. */
int wtg_mg_set_precision(wtg_mg* mg, int precision);
/* One level of forward_scalar_transform (wavelet.f90:104-150): on level
 * `fine`, wc_field := Compute_scalar_wavelets of s_field; on level fine - 1,
 * coarse_field := Restrict_scalar of (s_field, wc_field).  s_field must be
 * ghost filled (every field of this API is). */
int wtg_mg_forward(wtg_mg* mg, int fine, int s_field, int wc_field, int coarse_field);
/* One level of inverse_scalar_transform (wavelet.f90:298-346): on level
 * `fine`, fine_field := Prolong_scalar / Reconstruct_scalar of level
 * fine - 1's coarse_field and level fine's wc_field. */
int wtg_mg_inverse(wtg_mg* mg, int fine, int coarse_field, int wc_field, int fine_field);
/* The velocity restriction of forward_wavelet_transform (Restrict_velo,
 * wavelet.f90:1154-1174) on uniform levels: every coarse edge of coarse
 * slot := the mean of the two fine edges of fine slot that bisect it, all
 * layers (and members); then the coarse slot's ghost fill.  The levels
 * hold the same layers; coarse is the next coarser level.  Errors are
 * reported through wtg_last_error(fine). */
int wtg_restrict_velo(wtg_ctx* fine, int fine_slot, wtg_ctx* coarse, int coarse_slot);
/* Sub-sampling restriction (Restrict_ss, wavelet.f90:823-837, as for the
 * coarse Bernoulli and Exner functions): on level fine - 1, coarse_field :=
 * level fine's fine_field at the coincident nodes (poles included), halo
 * copies written in the same launch. */
int wtg_mg_restrict_ss(wtg_mg* mg, int fine, int fine_field, int coarse_field);

/* ---- multi-level explicit step on a full hierarchy ------------------------
 * One context per level, coarse to fine (consecutive levels, the same layers
 * and slots, FP64, single member), each holding its level's state in the
 * same slot.  On a full hierarchy (the reference's uniform refinement with
 * zero thresholds: every mask TOLRNZ) every coarse node has children, so the
 * coarse levels' Bernoulli and Exner functions, fluxes and velocity sources
 * are restrictions of the finer level's.  The contexts' streams are pointed
 * at the finest level's until wtg_ml_destroy; destroy the hierarchy before
 * its contexts. */
typedef struct wtg_ml wtg_ml;
/* Uploads the inter-level weights (overlap areas, their split parts, I_u_wgt,
 * R_F_wgt; wtg_grid_transfer_weights) and allocates the work buffers. */
int wtg_ml_create(const wtg_grid* const* grids, wtg_ctx* const* levels, int nlevels, wtg_ml** out);
void wtg_ml_destroy(wtg_ml* ml);
/* forward_wavelet_transform (wavelet.f90:20-101) of the slot: from the
 * finest level down, scalar wavelets, Restrict_scalar and Restrict_velo
 * into the coarser levels' slot, then the velocity wavelets of every level;
 * the wavelets stay in the hierarchy (wtg_ml_wavelets). */
int wtg_ml_forward(wtg_ml* ml, int slot);
/* inverse_wavelet_transform (wavelet.f90:208-296) of the slot from its
 * coarsest level and the hierarchy's wavelets: every finer level rebuilt. */
int wtg_ml_inverse(wtg_ml* ml, int slot);
/* Copies the wavelets of level (1..nlevels-1) to the host (to_device = 0) or
 * from it (1): mass, temp [L][ncells], velo [L][3][ncells]; null skips. */
int wtg_ml_wavelets(wtg_ml* ml, int level, int to_device, double* mass, double* temp, double* velo);
/* trend_ml (multi_level.f90:5-188) on every level: the finest level's
 * trend, the coarser levels' by compute-or-restrict (flux_cpt_restr,
 * scalar_cpt_restr, u_source_cpt_restr); the trend of each level goes to
 * dq_slot (owned cells, pole masters and owned edges). */
int wtg_ml_trend(wtg_ml* ml, int q_slot, int dq_slot);
/* WT_after_step (adapt.f90:192-237) of the slot: the wavelets of every
 * level from its state, then the inverse transform. */
int wtg_ml_wt_after_step(wtg_ml* ml, int slot);
/* RK4 (time_integr.f90:78-103) on the hierarchy: per stage trend_ml,
 * RK_sub_step on every level, WT_after_step.  Uses each context's two
 * stage buffers. */
int wtg_ml_rk4(wtg_ml* ml, int slot, double dt, int nsteps);
/* RK4_split (time_integr.f90:360-382, 444-478) on the hierarchy, the
 * reference's mode-split step of a multi-level ocean: per stage
 * barotropic_correction on every level, trend_ml, scalar_star / u_star with
 * each level's external pressure gradient (coarse free surfaces sub-sampled,
 * cpt_or_restr_eta), WT_after_step; then flux_divergence over levels (the
 * layer-integrated flux restricted by flux_cpt_restr), rhs_elliptic on every
 * level, the FMG free-surface solve on mg (the same contexts, wtg_mg_create,
 * with wtg_fs_set_params on every level; WTG_FS_ETA of each level is its free
 * surface), barotropic_correction, u_update and WT_after_step including the
 * free surface.  *iters: FMG V-cycles summed over levels and steps. */
int wtg_ml_rk4_split(wtg_ml* ml, wtg_mg* mg, int slot, int nsteps, const wtg_mg_params* p, int* iters);
/* RK3_split, RK2_split, Euler_split on the hierarchy (nstages 3, 2, 1; 4 is
 * wtg_ml_rk4_split) */
int wtg_ml_rk_split(wtg_ml* ml, wtg_mg* mg, int slot, int nstages, int nsteps, const wtg_mg_params* p, int* iters);

/* The finest level's 8x8 patches of owned nodes (n a multiple of 8): 10
 * (n/8)^2 of them, lozenge-major, then patch row, then patch column; nodes
 * row-major within a patch. */
int wtg_mg_npatches(const wtg_mg* mg);
/* Attaches (NULL detaches) the network of WTG_SOLVE_LEARNED; it must take a
 * batch of one row per patch.  The hierarchy does not own it: the network
 * must outlive the attachment (detach it, or destroy the hierarchy, before
 * wtg_mlp_destroy). */
typedef struct wtg_mlp wtg_mlp;
int wtg_mg_set_learned(wtg_mg* mg, wtg_mlp* mlp);
/* The network's input rows for the finest level's field `field`:
 * feats [npatch][192] = [r / sigma | (|topo| + eta_old) / 4000 m |
 * sin(latitude)] over the patch's 64 nodes, sigma [npatch] the RMS of r over
 * the patch; cells [npatch][64] (may be NULL) the cell of each patch node. */
int wtg_mg_features(wtg_mg* mg, int field, float* feats, double* sigma, int64_t* cells);

/* ---- learned correction: a 3-layer MLP on 8x8 node patches ---------------
 * No counterpart in the reference (its solvers are FMG and SRJ above).
 * y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3 per row, x [WTG_MLP_IN],
 * y [WTG_MLP_OUT]; weights and activations BF16, products accumulated in
 * FP32 on the matrix cores, FP32 master parameters.  Parameters flattened as
 * W1 [IN][HIDDEN], b1, W2 [HIDDEN][HIDDEN], b2, W3 [HIDDEN][OUT], b3
 * (WTG_MLP_NPARAMS floats).  One HIP stream per MLP; calls synchronise.
 * ------------------------------------------------------------------------- */
enum { WTG_MLP_IN = 192, WTG_MLP_HIDDEN = 128, WTG_MLP_OUT = 64, WTG_MLP_NPARAMS = 49472 };

/* He-normal weights from `seed` (the output layer's scaled by 0.1), zero
 * biases; activations for up to max_batch rows */
int wtg_mlp_create(int max_batch, uint64_t seed, wtg_mlp** out);
void wtg_mlp_destroy(wtg_mlp* mlp);
int wtg_mlp_last_error(const wtg_mlp* mlp, char* buf, size_t n);
int wtg_mlp_get_params(wtg_mlp* mlp, float* host, size_t n);
int wtg_mlp_set_params(wtg_mlp* mlp, const float* host, size_t n);  /* also resets the Adam state */
/* y [batch][OUT] = MLP(x [batch][IN]) (host arrays) */
int wtg_mlp_forward(wtg_mlp* mlp, const float* x, int batch, float* y);
/* nrep forward passes of the first `batch` rows already on the device (the
 * last wtg_mlp_forward's input), timed with HIP events: mean ms per pass */
int wtg_mlp_forward_timed(wtg_mlp* mlp, int batch, int nrep, double* ms_per_pass);
/* loss = sum (y - y_true)^2 / (2 batch) and its gradient with respect to the
 * parameters (WTG_MLP_NPARAMS floats), by the backward kernels */
int wtg_mlp_gradients(wtg_mlp* mlp, const float* x, const float* y_true, int batch, float* loss, float* grads);
/* training pairs resident on the device (n rows of x and y_true) */
int wtg_mlp_set_dataset(wtg_mlp* mlp, const float* x, const float* y_true, int64_t n);
/* nsteps Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on minibatches drawn
 * without replacement from shuffled epochs (seed); losses[nsteps] (may be
 * NULL) the minibatch losses before each step */
int wtg_mlp_train(wtg_mlp* mlp, int nsteps, int batch, float lr, uint64_t seed, float* losses);

#ifdef __cplusplus
}
#endif
#endif /* WAVETRISK_GPU_H */
