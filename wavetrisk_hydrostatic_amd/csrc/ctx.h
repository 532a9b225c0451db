// Internal: the device context shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wavetrisk_gpu.h"
#include "grid.h"

struct wtg_grid {
  wtg::Level lv;
};

// node fields of the free-surface (2-D) problem, one level
enum FsField {
  FS_ETA = 0,      // free-surface perturbation (sol(S_MASS, zlevels+1))
  FS_RHS,          // right-hand side of the elliptic equation (sol(S_TEMP, zlevels+1))
  FS_ETA_OLD,      // free surface of the previous step, in the operator (Laplacian_scalar(S_TEMP))
  FS_DIVF,         // flux divergence of the layer-integrated velocity (trend(S_MASS, zlevels+1))
  FS_DIVF_OLD,     // the same at the previous step (trend(S_TEMP, zlevels+1))
  FS_NUSER,        // fields above are reachable through the C ABI
  FS_W0 = FS_NUSER,  // solver work fields: W0..W0+7 BiCGSTAB, then the multigrid's
  MG_U2 = FS_W0 + 8,  // second buffer of the solution (Jacobi ping-pong)
  MG_RES,             // residual, two buffers
  MG_RES2,
  MG_CORR,            // V-cycle correction, two buffers
  MG_CORR2,
  MG_WC,              // scalar wavelet coefficients
  MG_PF,              // prolongation of the coarser correction
  FS_NFIELDS
};

// FP32 fields of the multigrid correction cycle (WTG_PRECISION_FP32_BULK)
enum Mg32Field {
  M32_CORR = 0,  // correction, two buffers
  M32_CORR2,
  M32_RES,       // residual of the correction equation, two buffers
  M32_RES2,
  M32_WC,        // scalar wavelet coefficients
  M32_PF,        // prolongation of the coarser correction
  M32_R,         // FP64 residual of the cycle's start, rounded (finest level)
  M32_NFIELDS
};

struct wtg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // stream of every launch (a hierarchy points it at its finest level's)
  hipStream_t own_stream = nullptr;  // the stream this context created
  wtg_grid_info info{};
  wtg_params params{};
  int nslots = 0;
  int nmembers = 1;       // ensemble members, each nlayers layers, one after another in every slot
  // set by a member-batched hierarchy solve on each of its levels: per
  // member 1 once that member's cycles have converged (device [nmembers])
  const double* mg_mdone = nullptr;
  bool split_fused = true;  // FP64 split-explicit stages: trend and AXPY in one launch
  int stage_version = 3;  // WTG_STAGE_KERNEL: 1 = tile loop, 2 = one thread per region cell, 3 = layer pairs
  bool topo_set = false;
  std::string err;
  // device geometry
  double *d_len = nullptr, *d_ped = nullptr, *d_pedf = nullptr, *d_area = nullptr, *d_hinv = nullptr;
  double *d_tri = nullptr, *d_cor = nullptr, *d_topo = nullptr;
  int64_t *d_nd = nullptr, *d_ed = nullptr, *d_fn = nullptr, *d_fe = nullptr;
  int8_t* d_es = nullptr;
  int nn = 0, ne = 0, nfn = 0, nfe = 0;
  std::vector<int64_t> node_src;
  // state slots: user slots then two stage buffers
  struct Slot {
    double *m = nullptr, *t = nullptr, *u = nullptr;
  };
  std::vector<Slot> slots;
  // precision of wtg_rk4 (WTG_PRECISION_*); the FP32-bulk mode's FP32
  // geometry copies and two FP32 stage-increment buffers are allocated on first use
  int precision = 0;
  bool f32_done = false;
  float *f_len = nullptr, *f_lp = nullptr, *f_pedf = nullptr, *f_area = nullptr, *f_hinv = nullptr;
  float* f_cor = nullptr;
  struct SlotF {
    float *m = nullptr, *t = nullptr, *u = nullptr;
  };
  SlotF fstage[2];
  int mixed_shape = WTG_SHAPE_LAYER_PAIRS;  // launch shape of the FP32-bulk stage kernel
  bool mixed_tuned = false;                 // false: timed on the first FP32-bulk step
  float tuned_ms[3] = {0.f, 0.f, 0.f};      // the autotuner's times of shapes 1..3
  // wtg_rk4 from a captured HIP graph (wtg_set_graph) and the key it was captured for
  bool use_graph = false;
  hipGraphExec_t rk_graph = nullptr;
  int g_slot = -1, g_precision = -1, g_shape = -1;
  double g_dt = 0.0;
  // free-surface problem
  double* d_lenf = nullptr;      // primal lengths of the flux (owner's for halo edges)
  int64_t* d_ncl = nullptr;      // node compute list: (dst, src) pairs, masters then ghost copies
  int nncl = 0, nmaster = 0;
  int64_t* d_rcell = nullptr;    // cells with a reduction weight
  double* d_rw = nullptr;        // their weights (visits of the reference's (0,1) sweep)
  int nred = 0;
  double fs_dt = 0, fs_theta1 = 1, fs_theta2 = 1;
  bool fs_params_set = false;
  double* fs[FS_NFIELDS] = {};   // allocated on first use
  float* fs32[M32_NFIELDS] = {}; // allocated on the first FP32 correction cycle
  double* d_part = nullptr;      // reduction partials
  unsigned* d_count = nullptr;   // blocks finished in a dot pass (the last one reduces)
  // dot passes of at most this many blocks finish in their last block (one
  // launch); larger ones use a final one-block kernel, because the last
  // block's device-scope fence costs L2 write-backs that grow with the grid
  int lastblock_max = 64;
  double* d_scal = nullptr;      // device scalars of the solvers
  double* h_scal = nullptr;      // pinned host copy (hipHostMalloc)
  int part_blocks = 0;
  int bicg_hint = 0;             // iterations of the last converged BiCGSTAB solve (its first chunk)
  double bicg_hint_tol = 0.0;    // and its tolerance
  // BiCGSTAB chunks of gated iterations replayed from a captured graph
  // (wtg_fs_set_graph), keyed by the solution and right-hand side arrays
  bool fs_graph = false;
  hipGraphExec_t bicg_graph = nullptr;
  const double* bicg_u = nullptr;
  const double* bicg_f = nullptr;
  double bicg_key[2] = {0.0, 0.0};  // operator coefficient and gravity of the captured graph
  const double* bicg_p = nullptr;   // buffers holding p and Ap when the graph was captured
  const double* bicg_Ap = nullptr;
  std::vector<void*> allocs;
  // ensembles: one single-member view per member for the free surface of
  // the split-explicit step (state, topography and free-surface fields at
  // the member's offset; geometry, lists and solver scalars shared, the
  // members' solves running one after another on the stream).  Owned here,
  // built on first use; a view owns no device memory.
  std::vector<wtg_ctx*> views;
  // Laplacian diffusion / hyperdiffusion of the trend (wtg_set_diffusion),
  // off while every order is 0; the pre-pass buffers are allocated for the
  // orders in use, [layer of every member][...][ncells]
  wtg_diffusion dif{};
  bool dif_on = false;
  double *dif_divu = nullptr, *dif_lapd = nullptr;  // div u and its Laplacian
  double *dif_lapm = nullptr, *dif_lapt = nullptr;  // Laplacians of mass and temperature
  double *dif_flux = nullptr;                       // [3] gradient fluxes of a Laplacian
  double *dif_vort = nullptr;                       // [2] vorticity (or rot rot) on LORT / UPLT
  double *dif_rotu = nullptr;                       // [3] rot u on edges
  double *dif_phm = nullptr, *dif_pht = nullptr;    // [3] scalar diffusive fluxes
  double *dif_phu = nullptr;                        // [3] velocity source
  // wtg_remap: a_vert then b_vert, [WTG_REMAP_MAX_LAYERS + 1] each (first use)
  double* d_remap_ab = nullptr;
  // vertical diffusion (wtg_set_vdiff): its constants, the eddy coefficients
  // at the interfaces [member][nlayers + 1][ncells], the TKE
  // [member][nlayers][ncells] and the forcing [member][6][ncells] (bottom and
  // top buoyancy flux, wind-stress magnitude, wind flux of the three edges)
  wtg_vdiff vd{};
  bool vd_set = false;
  double *vd_kt = nullptr, *vd_kv = nullptr, *vd_tke = nullptr, *vd_force = nullptr;
};

extern thread_local std::string g_create_error;
int wtg_fail(wtg_ctx* c, int code, const char* fmt, ...);

#define WTG_HIPCHK(ctx, call)                                                                         \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return wtg_fail(ctx, WTG_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

template <class Tp>
inline int wtg_dalloc(wtg_ctx* c, Tp** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(Tp) > 0 ? count * sizeof(Tp) : 1);
  if (e != hipSuccess)
    return wtg_fail(c, WTG_E_DEVICE, "hipMalloc(%zu bytes): %s", count * sizeof(Tp), hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = (Tp*)q;
  return 0;
}

// diffusion (diffusion.hip): buffers of the orders in c->dif, and the
// pre-pass that fills dif_phm / dif_pht / dif_phu from a ghost-filled state
int wtg_diffusion_alloc(wtg_ctx* c);
int wtg_diffusion_prepass(wtg_ctx* c, const double* m, const double* t, const double* u);

// allocates the free-surface fields on first use (elliptic.hip); an
// ensemble context holds them member-major ([nmembers][ncells] per field)
int wtg_fs_ready(wtg_ctx* c);
// flux_divergence into WTG_FS_DIVF and rhs_elliptic into WTG_FS_RHS of a
// single-member context (or view), queued without synchronising (elliptic.hip)
int wtg_fs_flux_div_async(wtg_ctx* c, int slot);
// BiCGSTAB on WTG_FS_ETA / WTG_FS_RHS of every member of a context at once
// (one launch per step of the iteration for all members, each member gated
// by its own convergence test); *iters, *err: the largest over the members
int wtg_fs_bicgstab_all(wtg_ctx* c, double tol, int maxit, int* iters, double* err);
int wtg_fs_rhs_async(wtg_ctx* c);

// split-explicit steps (trisk.hip) with the free-surface solve supplied by
// the caller: solve(user, c, iters) solves elliptic_lo(WTG_FS_ETA) =
// WTG_FS_RHS on c with WTG_FS_ETA_OLD set, the old eta as initial guess
typedef int (*wtg_fs_solver)(void* user, wtg_ctx* c, int* iters);
int wtg_split_steps(wtg_ctx* c, int slot, int nsteps, wtg_fs_solver solve, void* user, int* iters,
                    int nstages = 4);

// intermediate quantities of the finest level's stage that a hierarchy's
// coarser levels restrict (trisk.hip; null members are not kept)
struct WtgStageKeep {
  double *hm, *ht;     // [L][3][NC] mass / temperature fluxes on region cells [-1, n]^2
  double *bern, *exn;  // [L][NC] Bernoulli and Exner functions on [0, n]^2
  double* src;         // [L][3][NC] velocity source (-Qperp) on owned edges
  double *dm, *dt;     // [L][NC] scalar trends on owned cells and the pole masters
};
// one RK stage (or trend, trend_only = 1) of the FP64 layer-pair kernel
// with optional kept quantities
int wtg_launch_stage_keep(wtg_ctx* c, const wtg_ctx::Slot& y, const wtg_ctx::Slot& q, const wtg_ctx::Slot& out,
                          double h, int trend_only, const WtgStageKeep* keep);
// ghost fill of L layers of node fields m, t (skipped when m is null) and of
// an edge field u (skipped when null) with the context's lists
int wtg_ghost_fill_layers(wtg_ctx* c, int L, double* m, double* t, double* u);
// copies the never-filled halo slots of one state to another
int wtg_copy_frozen_slots(wtg_ctx* c, const wtg_ctx::Slot& from, const wtg_ctx::Slot& to);
// device address of the file-scope scratch of wtg_fs_dots_noalloc on a
// device, taken once per device (elliptic.hip); 0 on success
int wtg_noalloc_scratch_init(int device);
// split-explicit pieces on one level (trisk.hip): barotropic_correction of
// a state with the level's WTG_FS_ETA (masters only); out = base + h (trend
// + theta1 g grad eta_grad) (scalar_star, u_star); u -= theta1 dt g grad
// eta_grad (u_update)
int wtg_baro_correction_slot(wtg_ctx* c, const wtg_ctx::Slot& s);
int wtg_split_axpy(wtg_ctx* c, double h, const double* eta_grad, const wtg_ctx::Slot& base, const wtg_ctx::Slot& tr,
                   const wtg_ctx::Slot& out);
int wtg_u_update(wtg_ctx* c, const double* eta_grad, const wtg_ctx::Slot& s);
// flux_divergence of a slot into given buffers (elliptic.hip): the
// layer-integrated edge fluxes [3][ncells] on cells [-1, n]^2 and their
// divergence (compute list, halo copies included)
int wtg_fs_flux_div_into(wtg_ctx* c, int slot, double* flux3, double* divf);
// the free surface's own WT_after_step on a hierarchy (elliptic.hip): the
// scalar wavelets of WTG_FS_ETA on every level, then the inverse transform
// coarse to fine
struct wtg_mg;
int wtg_mg_eta_transform(wtg_mg* m);
// true when the solver hierarchy runs on exactly these contexts, coarse to fine
bool wtg_mg_same_levels(const wtg_mg* m, wtg_ctx* const* lv, int n);
