// MI355X kernels and C ABI of the free-surface elliptic problem (see the
// WTG_FS_* part of include/wavetrisk_gpu.h).
//
// Node fields live on the same lozenge tiles as the layer scalars, [ncells].
// Every kernel runs over the node compute list: one thread per master cell
// (owned nodes and the two pole masters), then one per ghost-filled halo
// cell, which evaluates the same expression at its owner and writes the
// halo copy.  So each output is ghost filled in the launch that writes it,
// with no separate copy kernel and no grid-wide synchronisation; kernels are
// out of place (a halo thread must read its owner's input, not its output).
//
// Arithmetic is FP64 in the reference's evaluation order, compiled without
// FMA contraction: the operator, its diagonal, a Jacobi sweep, the flux
// divergence and the right-hand side equal the CPU oracle bit for bit.  The
// norms and dot products reduce in a fixed tree order (deterministic run to
// run); they differ from a sequential sum in the last bits.
#include <algorithm>
#include <cmath>
#include <utility>

#include "ctx.h"

namespace {

constexpr int FS_NT = 256;        // threads per block
constexpr int RED_BLOCKS = 512;   // partial sums per dot product
constexpr int RED_MAX = 3;        // dot products fused in one pass

// device scalars of the solvers (d_scal)
enum Scal {
  SC_RHO = 0, SC_RHO_OLD, SC_ALPH, SC_OMGA, SC_B, SC_T0, SC_T1, SC_T2, SC_T3, SC_ERR, SC_NRMF, SC_ONE, SC_NSCAL
};

__device__ __forceinline__ double interp(double a, double b) { return 0.5 * (a + b); }

struct FsGeo {
  const double *lenf, *pedf, *hinv, *topo;
  int64_t NC;
  int S;
};

// flux of edge e of cell x (step1 itype 2, ops.f90:202-216):
//   g interp(|topo| + eta_old) * (difference of q along the edge) / len * pedlen
__device__ __forceinline__ double ell_flux(const FsGeo& g, const double* __restrict__ q,
                                           const double* __restrict__ eo, double grav, int64_t x, int e) {
  const int64_t nb = e == 0 ? x + 1 : (e == 1 ? x + g.S + 1 : x + g.S);
  const double cx = grav * (fabs(g.topo[x]) * 1.0 + eo[x]);
  const double cn = grav * (fabs(g.topo[nb]) * 1.0 + eo[nb]);
  const double d = e == 1 ? (q[x] / 1.0 - q[nb] / 1.0) : (q[nb] / 1.0 - q[x] / 1.0);
  return interp(cx, cn) * d / g.lenf[e * g.NC + x] * g.pedf[e * g.NC + x];
}

// elliptic_lo at cell c: coef * div(flux) - q  (div: ops.f90:1325-1343)
__device__ __forceinline__ double ell_apply_at(const FsGeo& g, const double* __restrict__ q,
                                               const double* __restrict__ eo, double grav, double coef, int64_t c) {
  const int64_t cW = c - 1, cSW = c - g.S - 1, cS = c - g.S;
  const double dv = (ell_flux(g, q, eo, grav, c, 0) - ell_flux(g, q, eo, grav, cW, 0) +
                     ell_flux(g, q, eo, grav, cSW, 1) - ell_flux(g, q, eo, grav, c, 1) +
                     ell_flux(g, q, eo, grav, c, 2) - ell_flux(g, q, eo, grav, cS, 2)) *
                    g.hinv[c];
  return coef * dv - q[c];
}

// elliptic_lo_diag at cell c (averaged depth, barotropic_2d.f90:286-322)
__device__ __forceinline__ double ell_diag_at(const FsGeo& g, const double* __restrict__ eo, double grav, double dt2,
                                              double th12, int64_t c) {
  const double depth = fabs(g.topo[c]) + eo[c] / 1.0;
  const double wgt = 2.0 * sqrt(3.0) * depth;
  const double lap = -grav * wgt * dt2 * g.hinv[c];
  return th12 * lap - 1.0;
}

__global__ void fs_apply_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, const double* __restrict__ q,
                                const double* __restrict__ eo, double grav, double coef, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[ncl[2 * t]] = ell_apply_at(g, q, eo, grav, coef, ncl[2 * t + 1]);
}

__global__ void fs_diag_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, const double* __restrict__ eo,
                               double grav, double dt2, double th12, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[ncl[2 * t]] = ell_diag_at(g, eo, grav, dt2, th12, ncl[2 * t + 1]);
}

// Jacobi_iteration (lin_solve.f90:301-340, cal_jacobi 342-359): out = u + w (f - L u) / diag
__global__ void fs_jacobi_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, const double* __restrict__ u,
                                 const double* __restrict__ f, const double* __restrict__ eo, double grav, double coef,
                                 double dt2, double th12, double w, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = ncl[2 * t + 1];
  const double Au = ell_apply_at(g, u, eo, grav, coef, c);
  const double dg = ell_diag_at(g, eo, grav, dt2, th12, c);
  out[ncl[2 * t]] = u[c] + w * (f[c] - Au) / dg;
}

// residual (lin_solve.f90:563-618): out = f - L u
__global__ void fs_residual_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, const double* __restrict__ u,
                                   const double* __restrict__ f, const double* __restrict__ eo, double grav,
                                   double coef, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = ncl[2 * t + 1];
  out[ncl[2 * t]] = f[c] - ell_apply_at(g, u, eo, grav, coef, c);
}

// lc (lin_solve.f90:520-561): out = a1 s1 + a2 s2, a = k (i < 0) or k * scal[i]
__global__ void fs_lc_kernel(int n, const int64_t* __restrict__ ncl, double k1, int i1, const double* __restrict__ s1,
                             double k2, int i2, const double* __restrict__ s2, const double* __restrict__ scal,
                             double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double a1 = i1 < 0 ? k1 : k1 * scal[i1];
  const double a2 = i2 < 0 ? k2 : k2 * scal[i2];
  const int64_t c = ncl[2 * t + 1];
  out[ncl[2 * t]] = a1 * s1[c] + a2 * s2[c];
}

// out = src (master and halo cells alike)
__global__ void fs_copy_kernel(int n, const int64_t* __restrict__ ncl, const double* __restrict__ s,
                               double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[ncl[2 * t]] = s[ncl[2 * t + 1]];
}

// flux_divergence (barotropic_2d.f90:326-366): div of sum_k u_k interp(mass_k) * pedlen
// (step1 itype 4, ops.f90:221-235); the mean state is zero
__device__ __forceinline__ double vflux(const FsGeo& g, int L, const double* __restrict__ m,
                                        const double* __restrict__ u, int64_t x, int e) {
  const int64_t nb = e == 0 ? x + 1 : (e == 1 ? x + g.S + 1 : x + g.S);
  double h = 0.0;
  for (int k = 0; k < L; ++k) {
    const double* mk = m + (int64_t)k * g.NC;
    const double dz0 = mk[x] + 0.0;
    const double dz = interp(dz0, mk[nb] + 0.0);
    h = h + u[(int64_t)(3 * k + e) * g.NC + x] * dz;
  }
  return h * g.pedf[e * g.NC + x];
}

__global__ void fs_flux_div_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, int L, const double* __restrict__ m,
                                   const double* __restrict__ u, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = ncl[2 * t + 1];
  const int64_t cW = c - 1, cSW = c - g.S - 1, cS = c - g.S;
  out[ncl[2 * t]] = (vflux(g, L, m, u, c, 0) - vflux(g, L, m, u, cW, 0) + vflux(g, L, m, u, cSW, 1) -
                     vflux(g, L, m, u, c, 1) + vflux(g, L, m, u, c, 2) - vflux(g, L, m, u, cS, 2)) *
                    g.hinv[c];
}

// rhs_elliptic (barotropic_2d.f90:168-184)
__global__ void fs_rhs_kernel(int n, const int64_t* __restrict__ ncl, const double* __restrict__ eta,
                              const double* __restrict__ divf, const double* __restrict__ divf_old, double dt,
                              double theta2, double rho0, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = ncl[2 * t + 1];
  out[ncl[2 * t]] = -eta[c] + dt * (theta2 * divf[c] + (1.0 - theta2) * divf_old[c]) / rho0;
}

// weighted dot products over the reduction cells: part[b*RED_MAX + k] =
// sum over this block's cells of w * (a_k * b_k); a fixed-order tree
__global__ __launch_bounds__(FS_NT) void fs_dot_kernel(int nred, const int64_t* __restrict__ rcell,
                                                       const double* __restrict__ rw, int nd,
                                                       const double* __restrict__ a0, const double* __restrict__ b0,
                                                       const double* __restrict__ a1, const double* __restrict__ b1,
                                                       const double* __restrict__ a2, const double* __restrict__ b2,
                                                       double* __restrict__ part) {
  __shared__ double sh[RED_MAX][FS_NT];
  const double* A[RED_MAX] = {a0, a1, a2};
  const double* B[RED_MAX] = {b0, b1, b2};
  double acc[RED_MAX] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * FS_NT + threadIdx.x; i < nred; i += gridDim.x * FS_NT) {
    const int64_t c = rcell[i];
    const double w = rw[i];
    for (int k = 0; k < RED_MAX; ++k)
      if (k < nd) acc[k] = acc[k] + w * (A[k][c] * B[k][c]);
  }
  for (int k = 0; k < RED_MAX; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = FS_NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < RED_MAX; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < RED_MAX) part[blockIdx.x * RED_MAX + threadIdx.x] = sh[threadIdx.x][0];
}

// scal[out_k] = sum of the partials of dot k (one block, fixed order)
__global__ __launch_bounds__(FS_NT) void fs_dot_final_kernel(int nblocks, int nd, const double* __restrict__ part,
                                                             int o0, int o1, int o2, double* __restrict__ scal) {
  __shared__ double sh[RED_MAX][FS_NT];
  double acc[RED_MAX] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += FS_NT)
    for (int k = 0; k < RED_MAX; ++k) acc[k] = acc[k] + part[b * RED_MAX + k];
  for (int k = 0; k < RED_MAX; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = FS_NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < RED_MAX; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int o[RED_MAX] = {o0, o1, o2};
    for (int k = 0; k < nd; ++k) scal[o[k]] = sh[k][0];
  }
}

// scalar steps of bicgstab (lin_solve.f90:395-422), in the reference's order
enum ScalOp { OP_ALPH = 0, OP_OMGA, OP_ERR, OP_BETA, OP_NRMF };
__global__ void fs_scalar_kernel(int op, double* __restrict__ s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  switch (op) {
    case OP_ALPH:  // alph = rho / dp(Ap, res0)
      s[SC_ALPH] = s[SC_RHO] / s[SC_T0];
      break;
    case OP_OMGA:  // omga = dp(As, s) / dp(As, As)
      s[SC_OMGA] = s[SC_T1] / s[SC_T2];
      break;
    case OP_ERR:  // err = l2(res) / nrm_f
      s[SC_ERR] = sqrt(s[SC_T0]) / s[SC_NRMF];
      break;
    case OP_BETA: {  // rho_old = rho; rho = dp(res0, res); b = (alph/omga) * (rho/rho_old)
      s[SC_RHO_OLD] = s[SC_RHO];
      s[SC_RHO] = s[SC_T3];
      s[SC_B] = (s[SC_ALPH] / s[SC_OMGA]) * (s[SC_RHO] / s[SC_RHO_OLD]);
      break;
    }
    case OP_NRMF: {  // nrm_f = l2(f), 1 if zero (FMG / SRJ, lin_solve.f90:73, 177)
      const double nf = sqrt(s[SC_T0]);
      s[SC_NRMF] = nf == 0.0 ? 1.0 : nf;
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline unsigned nblk(int n) { return (unsigned)((n + FS_NT - 1) / FS_NT); }

FsGeo fs_geo(const wtg_ctx* c) {
  return FsGeo{c->d_lenf, c->d_pedf, c->d_hinv, c->d_topo, c->info.ncells, c->info.S};
}

int fs_ready(wtg_ctx* c) {
  if (!c) return WTG_E_ARG;
  if (!c->topo_set) return wtg_fail(c, WTG_E_STATE, "wtg_set_topography must be called first");
  if (c->fs[0]) return 0;
  const int64_t NC = c->info.ncells;
  for (int f = 0; f < FS_NFIELDS; ++f) {
    if (int rc = wtg_dalloc<double>(c, &c->fs[f], NC)) return rc;
    WTG_HIPCHK(c, hipMemsetAsync(c->fs[f], 0, NC * sizeof(double), c->stream));
  }
  if (int rc = wtg_dalloc<double>(c, &c->d_part, (size_t)RED_BLOCKS * RED_MAX)) return rc;
  if (int rc = wtg_dalloc<double>(c, &c->d_scal, SC_NSCAL)) return rc;
  WTG_HIPCHK(c, hipMemsetAsync(c->d_scal, 0, SC_NSCAL * sizeof(double), c->stream));
  WTG_HIPCHK(c, hipHostMalloc((void**)&c->h_scal, SC_NSCAL * sizeof(double), hipHostMallocDefault));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int fs_field(wtg_ctx* c, int f, int nuser = FS_NUSER) {
  if (f < 0 || f >= nuser) return wtg_fail(c, WTG_E_ARG, "free-surface field %d out of range [0,%d)", f, nuser);
  return 0;
}

int fs_need_params(wtg_ctx* c) {
  if (!c->fs_params_set) return wtg_fail(c, WTG_E_STATE, "wtg_fs_set_params must be called first");
  return 0;
}

double fs_coef(const wtg_ctx* c) { return c->fs_theta1 * c->fs_theta2 * (c->fs_dt * c->fs_dt); }

int launch_apply(wtg_ctx* c, const double* q, double* out) {
  hipLaunchKernelGGL(fs_apply_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c), q,
                     c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c), out);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_residual(wtg_ctx* c, const double* u, const double* f, double* out) {
  hipLaunchKernelGGL(fs_residual_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c),
                     u, f, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c), out);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_jacobi(wtg_ctx* c, const double* u, const double* f, double w, double* out) {
  hipLaunchKernelGGL(fs_jacobi_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c), u,
                     f, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c), c->fs_dt * c->fs_dt,
                     c->fs_theta1 * c->fs_theta2, w, out);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_lc(wtg_ctx* c, double k1, int i1, const double* s1, double k2, int i2, const double* s2, double* out) {
  hipLaunchKernelGGL(fs_lc_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, k1, i1, s1, k2,
                     i2, s2, c->d_scal, out);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_copy(wtg_ctx* c, const double* s, double* out) {
  hipLaunchKernelGGL(fs_copy_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, s, out);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

// up to three dot products into scal[o0..]
int launch_dots(wtg_ctx* c, int nd, const double* a0, const double* b0, int o0, const double* a1 = nullptr,
                const double* b1 = nullptr, int o1 = 0, const double* a2 = nullptr, const double* b2 = nullptr,
                int o2 = 0) {
  const int nb = std::min(RED_BLOCKS, (int)nblk(c->nred));
  hipLaunchKernelGGL(fs_dot_kernel, dim3(nb), dim3(FS_NT), 0, c->stream, c->nred, c->d_rcell, c->d_rw, nd, a0, b0,
                     a1 ? a1 : a0, b1 ? b1 : b0, a2 ? a2 : a0, b2 ? b2 : b0, c->d_part);
  WTG_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(fs_dot_final_kernel, dim3(1), dim3(FS_NT), 0, c->stream, nb, nd, c->d_part, o0, o1, o2,
                     c->d_scal);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_scalar(wtg_ctx* c, int op) {
  hipLaunchKernelGGL(fs_scalar_kernel, dim3(1), dim3(64), 0, c->stream, op, c->d_scal);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int read_scalar(wtg_ctx* c, int i, double* out) {
  WTG_HIPCHK(c, hipMemcpyAsync(c->h_scal + i, c->d_scal + i, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = c->h_scal[i];
  return 0;
}

#define RC(x)                   \
  do {                          \
    if (int rc_ = (x)) return rc_; \
  } while (0)

// bicgstab (lin_solve.f90:361-431) on u with right-hand side f
int bicgstab(wtg_ctx* c, double* u, const double* f, double tol, int maxit, int* iters, double* err_out) {
  double** W = c->fs + FS_W0;
  double *res = W[0], *res0 = W[1], *p = W[2], *s = W[3], *Ap = W[4], *As = W[5], *corr = W[6], *ualt = W[7];
  double* ucur = u;
  RC(launch_dots(c, 1, f, f, SC_T0));  // nrm_f = l2(f)
  RC(launch_scalar(c, OP_NRMF));
  RC(launch_residual(c, ucur, f, res));
  RC(launch_copy(c, res, res0));
  RC(launch_copy(c, res0, p));
  RC(launch_copy(c, res0, s));
  RC(launch_apply(c, p, Ap));
  RC(launch_copy(c, Ap, As));
  RC(launch_dots(c, 1, res0, res, SC_RHO));
  int iter = 1;
  double err = 0.0;
  for (; iter <= maxit; ++iter) {
    RC(launch_dots(c, 1, Ap, res0, SC_T0));
    RC(launch_scalar(c, OP_ALPH));
    RC(launch_lc(c, 1.0, -1, res, -1.0, SC_ALPH, Ap, s));  // s = res - alph Ap
    RC(launch_apply(c, s, As));
    RC(launch_dots(c, 2, As, s, SC_T1, As, As, SC_T2));
    RC(launch_scalar(c, OP_OMGA));
    RC(launch_lc(c, 1.0, SC_ALPH, p, 1.0, SC_OMGA, s, corr));  // corr = alph p + omga s
    RC(launch_lc(c, 1.0, -1, ucur, 1.0, -1, corr, ualt));      // u = u + corr
    std::swap(ucur, ualt);
    RC(launch_lc(c, 1.0, -1, s, -1.0, SC_OMGA, As, res));  // res = s - omga As
    RC(launch_dots(c, 2, res, res, SC_T0, res0, res, SC_T3));
    RC(launch_scalar(c, OP_ERR));
    RC(read_scalar(c, SC_ERR, &err));
    if (err <= tol) break;
    RC(launch_scalar(c, OP_BETA));
    RC(launch_lc(c, 1.0, -1, p, -1.0, SC_OMGA, Ap, corr));  // corr = p - omga Ap
    RC(launch_lc(c, 1.0, -1, res, 1.0, SC_B, corr, p));     // p = res + b corr
    RC(launch_apply(c, p, Ap));
  }
  if (ucur != u) RC(launch_copy(c, ucur, u));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  *iters = iter;
  *err_out = err;
  return 0;
}

// Jacobi_iteration with weight 1 until the relative residual reaches tol
int jacobi_solve(wtg_ctx* c, double* u, const double* f, double tol, int maxit, int* iters, double* err_out) {
  double** W = c->fs + FS_W0;
  double *res = W[0], *ualt = W[7];
  double* ucur = u;
  RC(launch_dots(c, 1, f, f, SC_T0));
  RC(launch_scalar(c, OP_NRMF));
  int iter = 1;
  double err = 0.0;
  for (; iter <= maxit; ++iter) {
    RC(launch_jacobi(c, ucur, f, 1.0, ualt));
    std::swap(ucur, ualt);
    RC(launch_residual(c, ucur, f, res));
    RC(launch_dots(c, 1, res, res, SC_T0));
    RC(launch_scalar(c, OP_ERR));
    RC(read_scalar(c, SC_ERR, &err));
    if (err <= tol) break;
  }
  if (ucur != u) RC(launch_copy(c, ucur, u));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  *iters = iter;
  *err_out = err;
  return 0;
}

}  // namespace

extern "C" {

int wtg_fs_set_params(wtg_ctx* c, const wtg_fs_params* p) {
  if (!c || !p) return WTG_E_ARG;
  if (!(p->dt > 0.0)) return wtg_fail(c, WTG_E_ARG, "dt must be positive");
  c->fs_dt = p->dt;
  c->fs_theta1 = p->theta1;
  c->fs_theta2 = p->theta2;
  c->fs_params_set = true;
  return 0;
}

int wtg_fs_upload(wtg_ctx* c, int field, const double* host) {
  RC(fs_ready(c));
  RC(fs_field(c, field));
  if (!host) return wtg_fail(c, WTG_E_ARG, "null array");
  const int64_t NC = c->info.ncells;
  std::vector<double> t(host, host + NC);
  for (int64_t cc = 0; cc < NC; ++cc) {
    const int64_t s = c->node_src[cc];
    if (s >= 0) t[cc] = host[s];
  }
  WTG_HIPCHK(c, hipMemcpyAsync(c->fs[field], t.data(), NC * sizeof(double), hipMemcpyHostToDevice, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_fs_download(wtg_ctx* c, int field, double* host) {
  RC(fs_ready(c));
  RC(fs_field(c, field));
  if (!host) return wtg_fail(c, WTG_E_ARG, "null array");
  WTG_HIPCHK(c, hipMemcpyAsync(host, c->fs[field], c->info.ncells * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_elliptic_apply(wtg_ctx* c, int in_field, int out_field) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  RC(fs_field(c, in_field));
  RC(fs_field(c, out_field));
  if (in_field == out_field || out_field == FS_ETA_OLD)
    return wtg_fail(c, WTG_E_ARG, "output field must differ from the input and from WTG_FS_ETA_OLD");
  RC(launch_apply(c, c->fs[in_field], c->fs[out_field]));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_elliptic_diag(wtg_ctx* c, int out_field) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  RC(fs_field(c, out_field));
  if (out_field == FS_ETA_OLD) return wtg_fail(c, WTG_E_ARG, "output field must differ from WTG_FS_ETA_OLD");
  hipLaunchKernelGGL(fs_diag_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c),
                     c->fs[FS_ETA_OLD], c->params.grav, c->fs_dt * c->fs_dt, c->fs_theta1 * c->fs_theta2,
                     c->fs[out_field]);
  WTG_HIPCHK(c, hipGetLastError());
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_jacobi(wtg_ctx* c, int u_field, int f_field, double w, int nsweeps) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  RC(fs_field(c, u_field));
  RC(fs_field(c, f_field));
  if (u_field == f_field || u_field == FS_ETA_OLD) return wtg_fail(c, WTG_E_ARG, "u must differ from f and eta_old");
  if (nsweeps < 0) return wtg_fail(c, WTG_E_ARG, "nsweeps < 0");
  double* u = c->fs[u_field];
  double* alt = c->fs[FS_W0 + 7];
  for (int i = 0; i < nsweeps; ++i) {
    RC(launch_jacobi(c, u, c->fs[f_field], w, alt));
    std::swap(u, alt);
  }
  if (u != c->fs[u_field]) RC(launch_copy(c, u, c->fs[u_field]));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_fs_dot(wtg_ctx* c, int a_field, int b_field, double* out) {
  RC(fs_ready(c));
  RC(fs_field(c, a_field));
  RC(fs_field(c, b_field));
  if (!out) return wtg_fail(c, WTG_E_ARG, "null output");
  RC(launch_dots(c, 1, c->fs[a_field], c->fs[b_field], SC_T0));
  return read_scalar(c, SC_T0, out);
}

int wtg_flux_divergence(wtg_ctx* c, int slot) {
  RC(fs_ready(c));
  if (slot < 0 || slot >= c->nslots) return wtg_fail(c, WTG_E_ARG, "slot %d out of range [0,%d)", slot, c->nslots);
  const auto& s = c->slots[slot];
  hipLaunchKernelGGL(fs_flux_div_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c),
                     c->params.nlayers, s.m, s.u, c->fs[FS_DIVF]);
  WTG_HIPCHK(c, hipGetLastError());
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_rhs_elliptic(wtg_ctx* c) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  hipLaunchKernelGGL(fs_rhs_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, c->fs[FS_ETA],
                     c->fs[FS_DIVF], c->fs[FS_DIVF_OLD], c->fs_dt, c->fs_theta2, c->params.ref_density,
                     c->fs[FS_RHS]);
  WTG_HIPCHK(c, hipGetLastError());
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_elliptic_solve(wtg_ctx* c, int method, double rel_tol, int max_iter, int* iters, double* rel_resid) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  if (!iters || !rel_resid) return wtg_fail(c, WTG_E_ARG, "null output");
  if (max_iter < 0) return wtg_fail(c, WTG_E_ARG, "max_iter < 0");
  if (method == WTG_SOLVE_BICGSTAB) return bicgstab(c, c->fs[FS_ETA], c->fs[FS_RHS], rel_tol, max_iter, iters, rel_resid);
  if (method == WTG_SOLVE_JACOBI) return jacobi_solve(c, c->fs[FS_ETA], c->fs[FS_RHS], rel_tol, max_iter, iters, rel_resid);
  return wtg_fail(c, WTG_E_ARG, "unknown solver method %d", method);
}

}  // extern "C"
