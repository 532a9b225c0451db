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
#include <memory>
#include <utility>

#include "ctx.h"
#include "mg_ops.h"
#include "mlp.h"

namespace {

constexpr int FS_NT = 256;        // threads per block
constexpr int RED_BLOCKS = 512;   // partial sums per dot product
constexpr int RED_MAX = 3;        // dot products fused in one pass

// device scalars of the solvers (d_scal)
enum Scal {
  SC_RHO = 0, SC_RHO_OLD, SC_ALPH, SC_OMGA, SC_B, SC_T0, SC_T1, SC_T2, SC_T3, SC_ERR, SC_NRMF, SC_ONE,
  SC_DONE,  // 1 once an iterative solve has converged: gated kernels then do nothing
  SC_ITER,  // iterations completed by the gated loop
  SC_TOL,   // its tolerance
  SC_MR,    // minimal-residual step of the learned correction
  SC_NSCAL
};

// kernels of an iteration loop take the solver scalars as a gate and return
// at once when the loop has converged, so iterations can be queued ahead of
// the host's convergence test without changing the iterates
__device__ __forceinline__ bool gated(const double* __restrict__ gate) { return gate && gate[SC_DONE] != 0.0; }

// Batched solves of an ensemble (one launch for all members): member
// blockIdx.y, its node fields at blockIdx.y * ms (ms = cells per member),
// its solver scalars, dot partials and block counter in its own slices.
// One member: blockIdx.y = 0, no offset.
template <class Tp>
__device__ __forceinline__ Tp* mem(Tp* p, int64_t ms) {
  return p ? p + (int64_t)blockIdx.y * ms : p;
}

// Member-batched multigrid (a hierarchy of ensemble contexts, blockIdx.y =
// member): a member whose cycles have converged (mdone[member] != 0) keeps
// its fields -- kernels that write the next buffer of a solution copy it
// through, the others return; null mdone: no member has converged
__device__ __forceinline__ bool mdone_of(const double* __restrict__ mdone) {
  return mdone && mdone[blockIdx.y] != 0.0;
}

// scalar steps of the solvers (scalar_op below)
enum ScalOp { OP_ALPH = 0, OP_OMGA, OP_ERR, OP_BETA, OP_NRMF, OP_LOOP_INIT, OP_ERR_BETA, OP_MRSTEP, OP_NONE };

__device__ __forceinline__ double interp(double a, double b) { return 0.5 * (a + b); }

struct FsGeo {
  const double *lenf, *pedf, *hinv, *topo;
  int64_t NC;
  int S;
};

// flux of edge e of cell x (step1 itype 2, ops.f90:202-216):
//   g interp(|topo| + eta_old) * (difference of q along the edge) / len * pedlen
template <class F>
__device__ __forceinline__ double ell_flux(const FsGeo& g, const F* __restrict__ q,
                                           const double* __restrict__ eo, double grav, int64_t x, int e) {
  const int64_t nb = e == 0 ? x + 1 : (e == 1 ? x + g.S + 1 : x + g.S);
  const double cx = grav * (fabs(g.topo[x]) * 1.0 + eo[x]);
  const double cn = grav * (fabs(g.topo[nb]) * 1.0 + eo[nb]);
  const double d = e == 1 ? (q[x] / 1.0 - q[nb] / 1.0) : (q[nb] / 1.0 - q[x] / 1.0);
  return interp(cx, cn) * d / g.lenf[e * g.NC + x] * g.pedf[e * g.NC + x];
}

// elliptic_lo at cell c: coef * div(flux) - q  (div: ops.f90:1325-1343)
template <class F>
__device__ __forceinline__ double ell_apply_at(const FsGeo& g, const F* __restrict__ q,
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

__global__ void fs_apply_kernel(const double* __restrict__ gate, int n, const int64_t* __restrict__ ncl, FsGeo g,
                                const double* __restrict__ q, const double* __restrict__ eo, double grav, double coef,
                                double* __restrict__ out, double* __restrict__ out2, int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  gate = mem(gate, SC_NSCAL);
  if (t >= n || gated(gate)) return;
  g.topo = mem(g.topo, ms), q = mem(q, ms), eo = mem(eo, ms), out = mem(out, ms), out2 = mem(out2, ms);
  const double v = ell_apply_at(g, q, eo, grav, coef, ncl[2 * t + 1]);
  out[ncl[2 * t]] = v;
  if (out2) out2[ncl[2 * t]] = v;
}

// start of bicgstab (lin_solve.f90:371-388): res = f - L u, res0 = p = s =
// res; zero_u: the initial guess is 0 (u := 0 and L u = 0, u is not read)
__global__ void fs_bicg_init_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, double* __restrict__ u,
                                    const double* __restrict__ f, const double* __restrict__ eo, double grav,
                                    double coef, int zero_u, double* __restrict__ res, double* __restrict__ res0,
                                    double* __restrict__ p, double* __restrict__ s, int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  g.topo = mem(g.topo, ms), u = mem(u, ms), f = mem(f, ms), eo = mem(eo, ms);
  res = mem(res, ms), res0 = mem(res0, ms), p = mem(p, ms), s = mem(s, ms);
  const int64_t c = ncl[2 * t + 1], d = ncl[2 * t];
  const double r = zero_u ? f[c] - 0.0 : f[c] - ell_apply_at(g, (const double*)u, eo, grav, coef, c);
  res[d] = r;
  res0[d] = r;
  p[d] = r;
  s[d] = r;
  if (zero_u) u[d] = 0.0;
}

__global__ void fs_diag_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, const double* __restrict__ eo,
                               double grav, double dt2, double th12, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[ncl[2 * t]] = ell_diag_at(g, eo, grav, dt2, th12, ncl[2 * t + 1]);
}

// Jacobi_iteration (lin_solve.f90:301-340, cal_jacobi 342-359): out = u + w (f - L u) / diag
// (fields of type F: double, or float for the FP32 correction cycle; the
// arithmetic is FP64 either way)
// Zero: the sweep starts from u = 0 (L 0 = 0), u is not read
template <class F, bool Zero = false>
__global__ void fs_jacobi_kernel(const double* __restrict__ mdone, int n, const int64_t* __restrict__ ncl, FsGeo g,
                                 const F* __restrict__ u, const F* __restrict__ f, const double* __restrict__ eo,
                                 double grav, double coef, double dt2, double th12, double w, F* __restrict__ out,
                                 int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  g.topo = mem(g.topo, ms), u = mem(u, ms), f = mem(f, ms), eo = mem(eo, ms), out = mem(out, ms);
  const int64_t c = ncl[2 * t + 1];
  if (mdone_of(mdone)) {  // converged member: the sweep leaves u as it is
    out[ncl[2 * t]] = Zero ? (F)0.0 : u[c];
    return;
  }
  const double dg = ell_diag_at(g, eo, grav, dt2, th12, c);
  if (Zero) {
    out[ncl[2 * t]] = (F)(0.0 + w * (f[c] - 0.0) / dg);
    return;
  }
  const double Au = ell_apply_at(g, u, eo, grav, coef, c);
  out[ncl[2 * t]] = (F)(u[c] + w * (f[c] - Au) / dg);
}

// residual (lin_solve.f90:563-618): out = f - L u
template <class Fu, class Ff, class Fo>
__global__ void fs_residual_kernel(const double* __restrict__ mdone, int n, const int64_t* __restrict__ ncl, FsGeo g,
                                   const Fu* __restrict__ u, const Ff* __restrict__ f,
                                   const double* __restrict__ eo, double grav, double coef, Fo* __restrict__ out,
                                   int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || mdone_of(mdone)) return;
  g.topo = mem(g.topo, ms), u = mem(u, ms), f = mem(f, ms), eo = mem(eo, ms), out = mem(out, ms);
  const int64_t c = ncl[2 * t + 1];
  out[ncl[2 * t]] = (Fo)(f[c] - ell_apply_at(g, u, eo, grav, coef, c));
}

// lc (lin_solve.f90:520-561): out = a1 s1 + a2 s2, a = k (i < 0) or k * scal[i]
template <class F1, class F2, class Fo>
__global__ void fs_lc_kernel(const double* __restrict__ gate, const double* __restrict__ mdone, int n,
                             const int64_t* __restrict__ ncl, double k1, int i1, const F1* __restrict__ s1, double k2,
                             int i2, const F2* __restrict__ s2, const double* __restrict__ scal, Fo* __restrict__ out,
                             int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  gate = mem(gate, SC_NSCAL);
  if (t >= n || gated(gate)) return;
  scal = mem(scal, SC_NSCAL), s1 = mem(s1, ms), s2 = mem(s2, ms), out = mem(out, ms);
  const int64_t c = ncl[2 * t + 1];
  if (mdone_of(mdone)) {  // converged member: out is its first operand (u of u + a e)
    out[ncl[2 * t]] = (Fo)s1[c];
    return;
  }
  const double a1 = i1 < 0 ? k1 : k1 * scal[i1];
  const double a2 = i2 < 0 ? k2 : k2 * scal[i2];
  out[ncl[2 * t]] = (Fo)(a1 * s1[c] + a2 * s2[c]);
}

// halo copies of a field from its masters (entries nmaster.. of the list)
__global__ void fs_ghost_copy_kernel(int nmaster, int n, const int64_t* __restrict__ ncl, double* __restrict__ f,
                                     int64_t ms) {
  const int t = nmaster + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  f = mem(f, ms);
  f[ncl[2 * t]] = f[ncl[2 * t + 1]];
}

// out = src (master and halo cells alike), converted to the output type
template <class Fi, class Fo>
__global__ void fs_copy_kernel(int n, const int64_t* __restrict__ ncl, const Fi* __restrict__ s,
                               Fo* __restrict__ out, int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  s = mem(s, ms), out = mem(out, ms);
  out[ncl[2 * t]] = (Fo)s[ncl[2 * t + 1]];
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

// The divergence in two passes, each edge's layer sum evaluated once:
// fs_edge_flux_kernel writes vflux(x, e) for every cell x = (lz, i, j) with
// i, j in [-1, n] (the masters, poles included, and their W, SW, S
// neighbours; every neighbour read stays in the lozenge's tile), then
// fs_div_kernel sums (F0(c) - F0(W) + F1(SW) - F1(c) + F2(c) - F2(S)) hex_inv
// in the order of div (ops.f90:1325-1343).
__global__ void fs_edge_flux_kernel(int nlz, FsGeo g, int L, const double* __restrict__ m,
                                    const double* __restrict__ u, double* __restrict__ F0, double* __restrict__ F1,
                                    double* __restrict__ F2) {
  const int w = nlz + 2;  // cells per row and rows per lozenge in the pass
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)10 * w * w) return;
  const int lz = (int)(t / ((int64_t)w * w)), r = (int)(t % ((int64_t)w * w));
  const int i = r % w - 1, j = r / w - 1;
  const int64_t x = (int64_t)lz * g.S * g.S + (int64_t)(j + 2) * g.S + (i + 2);
  F0[x] = vflux(g, L, m, u, x, 0);
  F1[x] = vflux(g, L, m, u, x, 1);
  F2[x] = vflux(g, L, m, u, x, 2);
}

__global__ void fs_div_kernel(int n, const int64_t* __restrict__ ncl, FsGeo g, const double* __restrict__ F0,
                              const double* __restrict__ F1, const double* __restrict__ F2,
                              double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = ncl[2 * t + 1];
  const int64_t cW = c - 1, cSW = c - g.S - 1, cS = c - g.S;
  out[ncl[2 * t]] = (F0[c] - F0[cW] + F1[cSW] - F1[c] + F2[c] - F2[cS]) * g.hinv[c];
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
// Sum over the block of acc[k] into sh[k][0] (read by thread 0): pairwise
// tree, thread t adding t + s for s = FS_NT/2 .. 1.  The steps across waves
// go through LDS; the last six (s = 32 .. 1) run in wave 0 with shuffles,
// which pair the same lanes, so the sums are those of the all-LDS tree bit
// for bit, without their barriers.
__device__ __forceinline__ void block_tree(double acc[RED_MAX], double (*sh)[FS_NT]) {
  static_assert(FS_NT >= 128 && FS_NT % 64 == 0, "tree over whole waves");
  for (int k = 0; k < RED_MAX; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = FS_NT / 2; s >= 64; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < RED_MAX; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 64) {
    double v[RED_MAX];
    for (int k = 0; k < RED_MAX; ++k) v[k] = sh[k][threadIdx.x];
    for (int s = 32; s > 0; s >>= 1)
      for (int k = 0; k < RED_MAX; ++k) v[k] = v[k] + __shfl_down(v[k], s, 64);
    if (threadIdx.x == 0)
      for (int k = 0; k < RED_MAX; ++k) sh[k][0] = v[k];
  }
}

__global__ __launch_bounds__(FS_NT) void fs_dot_kernel(const double* __restrict__ gate, int nred,
                                                       const int64_t* __restrict__ rcell,
                                                       const double* __restrict__ rw, int nd,
                                                       const double* __restrict__ a0, const double* __restrict__ b0,
                                                       const double* __restrict__ a1, const double* __restrict__ b1,
                                                       const double* __restrict__ a2, const double* __restrict__ b2,
                                                       double* __restrict__ part, int64_t ms) {
  gate = mem(gate, SC_NSCAL);
  if (gated(gate)) return;
  a0 = mem(a0, ms), b0 = mem(b0, ms), a1 = mem(a1, ms), b1 = mem(b1, ms), a2 = mem(a2, ms), b2 = mem(b2, ms);
  part = mem(part, (int64_t)RED_BLOCKS * RED_MAX);
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
  block_tree(acc, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < RED_MAX; ++k) part[blockIdx.x * RED_MAX + k] = sh[k][0];
}

// scal[out_k] = sum of the partials of dot k (one block, fixed order)
__device__ void scalar_op(int op, int gated_loop, double arg, double* __restrict__ s);

// ... and then the scalar step post_op of the solver (OP_NONE for none)
__global__ __launch_bounds__(FS_NT) void fs_dot_final_kernel(int gated_loop, int nblocks, int nd,
                                                             const double* __restrict__ part, int o0, int o1, int o2,
                                                             double* __restrict__ scal, int post_op) {
  scal = mem(scal, SC_NSCAL);
  part = mem(part, (int64_t)RED_BLOCKS * RED_MAX);
  if (gated_loop && scal[SC_DONE] != 0.0) return;
  __shared__ double sh[RED_MAX][FS_NT];
  double acc[RED_MAX] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += FS_NT)
    for (int k = 0; k < RED_MAX; ++k) acc[k] = acc[k] + part[b * RED_MAX + k];
  block_tree(acc, sh);
  if (threadIdx.x == 0) {
    const int o[RED_MAX] = {o0, o1, o2};
    for (int k = 0; k < nd; ++k) scal[o[k]] = sh[k][0];
    if (post_op != OP_NONE) scalar_op(post_op, gated_loop, 0.0, scal);
  }
}

// scalar steps of bicgstab (lin_solve.f90:395-422), in the reference's order
// gated_loop: the op belongs to an iteration of a gated loop (no-op once it
// has converged; OP_ERR then also counts the iteration and tests it against
// SC_TOL).  OP_ERR_BETA is OP_ERR followed, if not converged, by OP_BETA.
__device__ void scalar_op(int op, int gated_loop, double arg, double* __restrict__ s) {
  if (gated_loop && s[SC_DONE] != 0.0) return;
  switch (op) {
    case OP_LOOP_INIT:
      s[SC_DONE] = 0.0;
      s[SC_ITER] = 0.0;
      s[SC_TOL] = arg;
      break;
    case OP_ALPH:  // alph = rho / dp(Ap, res0)
      s[SC_ALPH] = s[SC_RHO] / s[SC_T0];
      break;
    case OP_OMGA:  // omga = dp(As, s) / dp(As, As)
      s[SC_OMGA] = s[SC_T1] / s[SC_T2];
      break;
    case OP_ERR:
    case OP_ERR_BETA:  // err = l2(res) / nrm_f
      s[SC_ERR] = sqrt(s[SC_T0]) / s[SC_NRMF];
      if (gated_loop) {
        s[SC_ITER] = s[SC_ITER] + 1.0;
        if (s[SC_ERR] <= s[SC_TOL]) {
          s[SC_DONE] = 1.0;
          break;
        }
      }
      if (op == OP_ERR) break;
      [[fallthrough]];
    case OP_BETA: {  // rho_old = rho; rho = dp(res0, res); b = (alph/omga) * (rho/rho_old)
      s[SC_RHO_OLD] = s[SC_RHO];
      s[SC_RHO] = s[SC_T3];
      s[SC_B] = (s[SC_ALPH] / s[SC_OMGA]) * (s[SC_RHO] / s[SC_RHO_OLD]);
      break;
    }
    case OP_MRSTEP: {  // a = <r, L e> / <L e, L e> if that is positive and finite, else 0
      const double a = s[SC_T1] > 0.0 ? s[SC_T0] / s[SC_T1] : 0.0;
      s[SC_MR] = (a > 0.0 && isfinite(a)) ? a : 0.0;
      break;
    }
    case OP_NRMF: {  // nrm_f = l2(f), 1 if zero (FMG / SRJ, lin_solve.f90:73, 177)
      const double nf = sqrt(s[SC_T0]);
      s[SC_NRMF] = nf == 0.0 ? 1.0 : nf;
      break;
    }
    default:
      break;
  }
}

__global__ void fs_scalar_kernel(int op, int gated_loop, double arg, double* __restrict__ s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  scalar_op(op, gated_loop, arg, mem(s, SC_NSCAL));
}


// ---------------------------------------------------------------------------
// One-launch dot products and the fused BiCGSTAB steps
// ---------------------------------------------------------------------------
// A dot pass writes one partial sum per block (fs_dot_kernel's order), and
// the block that finishes last -- counted with an atomic, no block waits --
// reduces the partials in fs_dot_final_kernel's order and runs the scalar
// step.  The partials are read with device-scope loads (they were written by
// other compute units).  The counter is reset by that last block.

// tree over this block's RED_MAX accumulators; thread k < RED_MAX writes part[b][k]
__device__ __forceinline__ void dot_block_tree(double acc[RED_MAX], double (*sh)[FS_NT], double* part, int b) {
  block_tree(acc, sh);
  if (part && threadIdx.x == 0)
    for (int k = 0; k < RED_MAX; ++k) part[b * RED_MAX + k] = sh[k][0];
}

__device__ __forceinline__ void dot_last_block(int nbred, int nd, const double* part, int o0, int o1, int o2,
                                               double* __restrict__ scal, int post_op, int gated_loop,
                                               unsigned* count, double (*sh)[FS_NT]) {
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(count, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  double acc[RED_MAX] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nbred; b += FS_NT)
    for (int k = 0; k < RED_MAX; ++k)
      acc[k] = acc[k] + __hip_atomic_load(part + b * RED_MAX + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  dot_block_tree(acc, sh, nullptr, 0);
  if (threadIdx.x == 0) {
    const int o[RED_MAX] = {o0, o1, o2};
    for (int k = 0; k < nd; ++k) scal[o[k]] = sh[k][0];
    if (post_op != OP_NONE) scalar_op(post_op, gated_loop, 0.0, scal);
    atomicExch(count, 0u);
  }
}

// weighted dot products into scal[o_k], then post_op (one launch)
__global__ __launch_bounds__(FS_NT) void fs_dot1_kernel(const double* __restrict__ gate, int gated_loop, int nred,
                                                        const int64_t* __restrict__ rcell,
                                                        const double* __restrict__ rw, int nd,
                                                        const double* __restrict__ a0, const double* __restrict__ b0,
                                                        const double* __restrict__ a1, const double* __restrict__ b1,
                                                        const double* __restrict__ a2, const double* __restrict__ b2,
                                                        double* __restrict__ part, int o0, int o1, int o2,
                                                        double* __restrict__ scal, int post_op, unsigned* count,
                                                        int64_t ms) {
  gate = mem(gate, SC_NSCAL);
  if (gated(gate)) return;
  a0 = mem(a0, ms), b0 = mem(b0, ms), a1 = mem(a1, ms), b1 = mem(b1, ms), a2 = mem(a2, ms), b2 = mem(b2, ms);
  part = mem(part, (int64_t)RED_BLOCKS * RED_MAX), scal = mem(scal, SC_NSCAL), count = mem(count, 1);
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
  dot_block_tree(acc, sh, part, blockIdx.x);
  if (count) dot_last_block(gridDim.x, nd, part, o0, o1, o2, scal, post_op, gated_loop, count, sh);
}

// elliptic_lo at cell c from the values q7 of the field at c and its six
// neighbours {c, E, NE, N, W, SW, S}: the expression of ell_apply_at
__device__ __forceinline__ double ell_flux_v(const FsGeo& g, const double* __restrict__ eo, double grav, int64_t x,
                                             int e, double qx, double qn) {
  const int64_t nb = e == 0 ? x + 1 : (e == 1 ? x + g.S + 1 : x + g.S);
  const double cx = grav * (fabs(g.topo[x]) * 1.0 + eo[x]);
  const double cn = grav * (fabs(g.topo[nb]) * 1.0 + eo[nb]);
  const double d = e == 1 ? (qx / 1.0 - qn / 1.0) : (qn / 1.0 - qx / 1.0);
  return interp(cx, cn) * d / g.lenf[e * g.NC + x] * g.pedf[e * g.NC + x];
}
__device__ __forceinline__ double ell_apply_v(const FsGeo& g, const double* __restrict__ eo, double grav,
                                              double coef, int64_t c, const double q7[7]) {
  const int64_t cW = c - 1, cSW = c - g.S - 1, cS = c - g.S;
  const double dv = (ell_flux_v(g, eo, grav, c, 0, q7[0], q7[1]) - ell_flux_v(g, eo, grav, cW, 0, q7[4], q7[0]) +
                     ell_flux_v(g, eo, grav, cSW, 1, q7[5], q7[0]) - ell_flux_v(g, eo, grav, c, 1, q7[0], q7[2]) +
                     ell_flux_v(g, eo, grav, c, 2, q7[0], q7[3]) - ell_flux_v(g, eo, grav, cS, 2, q7[6], q7[0])) *
                    g.hinv[c];
  return coef * dv - q7[0];
}
__device__ __forceinline__ void cells7(const FsGeo& g, int64_t c, int64_t x7[7]) {
  x7[0] = c;
  x7[1] = c + 1;
  x7[2] = c + g.S + 1;
  x7[3] = c + g.S;
  x7[4] = c - 1;
  x7[5] = c - g.S - 1;
  x7[6] = c - g.S;
}

// s = res - alph Ap and As = L s in one pass (s is evaluated at the stencil's
// seven cells; res and Ap are ghost filled, so halo values equal the owners')
__global__ void fs_bicg_s_apply_kernel(const double* __restrict__ gate, int n, const int64_t* __restrict__ ncl,
                                       FsGeo g, const double* __restrict__ res, const double* __restrict__ Ap,
                                       const double* __restrict__ eo, double grav, double coef,
                                       const double* __restrict__ scal, double* __restrict__ s_out,
                                       double* __restrict__ As_out, int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  gate = mem(gate, SC_NSCAL);
  if (t >= n || gated(gate)) return;
  g.topo = mem(g.topo, ms), res = mem(res, ms), Ap = mem(Ap, ms), eo = mem(eo, ms);
  scal = mem(scal, SC_NSCAL), s_out = mem(s_out, ms), As_out = mem(As_out, ms);
  const double a2 = -1.0 * scal[SC_ALPH];
  const int64_t c = ncl[2 * t + 1], dst = ncl[2 * t];
  int64_t x7[7];
  cells7(g, c, x7);
  double q7[7];
  for (int k = 0; k < 7; ++k) q7[k] = 1.0 * res[x7[k]] + a2 * Ap[x7[k]];
  s_out[dst] = q7[0];
  As_out[dst] = ell_apply_v(g, eo, grav, coef, c, q7);
}

// p_new = res + b (p - omga Ap) and Ap_new = L p_new in one pass
__global__ void fs_bicg_p_apply_kernel(const double* __restrict__ gate, int n, const int64_t* __restrict__ ncl,
                                       FsGeo g, const double* __restrict__ res, const double* __restrict__ p,
                                       const double* __restrict__ Ap, const double* __restrict__ eo, double grav,
                                       double coef, const double* __restrict__ scal, double* __restrict__ pn_out,
                                       double* __restrict__ Apn_out, int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  gate = mem(gate, SC_NSCAL);
  if (t >= n || gated(gate)) return;
  g.topo = mem(g.topo, ms), res = mem(res, ms), p = mem(p, ms), Ap = mem(Ap, ms), eo = mem(eo, ms);
  scal = mem(scal, SC_NSCAL), pn_out = mem(pn_out, ms), Apn_out = mem(Apn_out, ms);
  const double bb = 1.0 * scal[SC_B], mo = -1.0 * scal[SC_OMGA];
  const int64_t c = ncl[2 * t + 1], dst = ncl[2 * t];
  int64_t x7[7];
  cells7(g, c, x7);
  double q7[7];
  for (int k = 0; k < 7; ++k) {
    const double corr = 1.0 * p[x7[k]] + mo * Ap[x7[k]];
    q7[k] = 1.0 * res[x7[k]] + bb * corr;
  }
  pn_out[dst] = q7[0];
  Apn_out[dst] = ell_apply_v(g, eo, grav, coef, c, q7);
}

// u = u + (alph p + omga s) on the masters, res = s - omga As, and the dots
// (res, res) -> T0, (res0, res) -> T3 with err, the convergence test and the
// next rho and b (OP_ERR_BETA) in one launch; res is evaluated where the dots
// read it.  Blocks b < nbred also reduce (the partition of fs_dot1_kernel).
__global__ __launch_bounds__(FS_NT) void fs_bicg_res_kernel(
    const double* __restrict__ gate, int nmaster, int nncl, const int64_t* __restrict__ ncl, int nred, int nbred,
    const int64_t* __restrict__ rcell, const double* __restrict__ rw, double* __restrict__ u,
    const double* __restrict__ p, const double* __restrict__ s, const double* __restrict__ As,
    const double* __restrict__ res0, double* __restrict__ scal, double* __restrict__ res_out,
    double* __restrict__ part, unsigned* count, int64_t ms) {
  gate = mem(gate, SC_NSCAL);
  if (gated(gate)) return;
  u = mem(u, ms), p = mem(p, ms), s = mem(s, ms), As = mem(As, ms), res0 = mem(res0, ms), res_out = mem(res_out, ms);
  scal = mem(scal, SC_NSCAL), part = mem(part, (int64_t)RED_BLOCKS * RED_MAX), count = mem(count, 1);
  __shared__ double sh[RED_MAX][FS_NT];
  const double alph = scal[SC_ALPH], omga = scal[SC_OMGA];
  const int i = blockIdx.x * FS_NT + threadIdx.x;
  if (i < nmaster) {
    const int64_t c = ncl[2 * i + 1];
    const double corr = (1.0 * alph) * p[c] + (1.0 * omga) * s[c];
    u[c] = 1.0 * u[c] + 1.0 * corr;
  }
  const double mo = -1.0 * omga;
  if (i < nncl) {
    const int64_t c = ncl[2 * i + 1];
    res_out[ncl[2 * i]] = 1.0 * s[c] + mo * As[c];
  }
  double acc[RED_MAX] = {0.0, 0.0, 0.0};
  if ((int)blockIdx.x < nbred)
    for (int j = blockIdx.x * FS_NT + threadIdx.x; j < nred; j += nbred * FS_NT) {
      const int64_t c = rcell[j];
      const double w = rw[j];
      const double r = 1.0 * s[c] + mo * As[c];
      acc[0] = acc[0] + w * (r * r);
      acc[1] = acc[1] + w * (res0[c] * r);
    }
  dot_block_tree(acc, sh, (int)blockIdx.x < nbred ? part : nullptr, blockIdx.x);
  if (count) dot_last_block(nbred, 2, part, SC_T0, SC_T3, 0, scal, OP_ERR_BETA, 1, count, sh);
}

// ===========================================================================
// Multi-level scalar wavelet transform (one level) and the multigrid solvers
// ===========================================================================
// (MgGeo and the point stencils of the transform: mg_ops.h)

// Compute_scalar_wavelets (wavelet.f90:648-670) over the fine compute list:
// wc = s - Interp_node at midpoints, 0 at coincident nodes
template <class F>
__global__ void mg_wavelet_kernel(const double* __restrict__ mdone, int n, const int64_t* __restrict__ ncl, MgGeo g,
                                  const F* __restrict__ s, F* __restrict__ wc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || mdone_of(mdone)) return;
  s = mem(s, g.NCf), wc = mem(wc, g.NCf);
  const int64_t x = ncl[2 * t + 1];
  int lz, i, j;
  mg_ij(x, g.Sf, lz, i, j);
  wc[ncl[2 * t]] = (F)(((i | j) & 1) ? s[x] - mg_interp(g, s, x, i, j) : 0.0);
}

// Restrict_scalar over the coarse compute list: s_c = s_f(child) + lift * hex_inv_c
template <class F>
__global__ void mg_restrict_kernel(const double* __restrict__ mdone, int n, const int64_t* __restrict__ ncl, MgGeo g,
                                   const F* __restrict__ sf, const F* __restrict__ wc, F* __restrict__ sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || mdone_of(mdone)) return;
  sf = mem(sf, g.NCf), wc = mem(wc, g.NCf), sc = mem(sc, g.NCc);
  const int64_t p = ncl[2 * t + 1];
  int lz, i, j;
  mg_ij(p, g.Sc, lz, i, j);
  const int64_t c = (int64_t)lz * g.Sf * g.Sf + (int64_t)(2 * j + 2) * g.Sf + (2 * i + 2);
  const bool pole = p == g.sp_c || p == g.np_c;
  const double lift = pole ? mg_pole_lift(g, wc, p == g.sp_c, false) : mg_lift(g, wc, c, true);
  sc[ncl[2 * t]] = (F)(sf[c] + lift * g.hinv_c[p]);
}

// Restrict_ss (wavelet.f90:823-837) over the coarse compute list: the
// coarse value is the fine value at the coincident child (poles included)
__global__ void mg_restrict_ss_kernel(int n, const int64_t* __restrict__ ncl, MgGeo g, const double* __restrict__ sf,
                                      double* __restrict__ sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t p = ncl[2 * t + 1];
  int lz, i, j;
  mg_ij(p, g.Sc, lz, i, j);
  sc[ncl[2 * t]] = sf[(int64_t)lz * g.Sf * g.Sf + (int64_t)(2 * j + 2) * g.Sf + (2 * i + 2)];
}

// Prolong_scalar over the fine compute list, coincident nodes:
// s_f = s_c(parent) - lift * hex_inv_c
// ZeroW: all wavelets zero (prol_fun), wc is not read
template <class F, bool ZeroW = false>
__global__ void mg_prolong_even_kernel(const double* __restrict__ mdone, int n, const int64_t* __restrict__ ncl,
                                       MgGeo g, const F* __restrict__ sc, const F* __restrict__ wc,
                                       F* __restrict__ sf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || mdone_of(mdone)) return;
  sc = mem(sc, g.NCc), wc = mem(wc, g.NCf), sf = mem(sf, g.NCf);
  const int64_t x = ncl[2 * t + 1];
  int lz, i, j;
  mg_ij(x, g.Sf, lz, i, j);
  if ((i | j) & 1) return;
  const int64_t p = (int64_t)lz * g.Sc * g.Sc + (int64_t)(j / 2 + 2) * g.Sc + (i / 2 + 2);
  if (ZeroW) {
    sf[ncl[2 * t]] = sc[p];
    return;
  }
  const bool pole = x == g.sp_f || x == g.np_f;
  const double lift = pole ? mg_pole_lift(g, wc, x == g.sp_f, true) : mg_lift(g, wc, x, false);
  sf[ncl[2 * t]] = (F)(sc[p] - lift * g.hinv_c[p]);
}

// Reconstruct_scalar (wavelet.f90:613-646) over the fine compute list,
// midpoints: s_f = Interp_node(coincident s_f) + wc.  Reads only coincident
// cells and writes only midpoints, so it runs in place after the even pass.
template <class F, bool ZeroW = false>
__global__ void mg_prolong_odd_kernel(const double* __restrict__ mdone, int n, const int64_t* __restrict__ ncl,
                                      MgGeo g, const F* __restrict__ wc, F* __restrict__ sf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || mdone_of(mdone)) return;
  wc = mem(wc, g.NCf), sf = mem(sf, g.NCf);
  const int64_t x = ncl[2 * t + 1];
  int lz, i, j;
  mg_ij(x, g.Sf, lz, i, j);
  if (!((i | j) & 1)) return;
  sf[ncl[2 * t]] = ZeroW ? (F)mg_interp(g, sf, x, i, j) : (F)(mg_interp(g, sf, x, i, j) + wc[x]);
}

// member-batched cycles: the converged members' BiCGSTAB starts converged
// (SC_DONE), so its gated iterations leave their coarse fields alone
__global__ void mg_mark_done_kernel(int M, const double* __restrict__ mdone, double* __restrict__ scal) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < M && mdone[k] != 0.0) scal[(int64_t)k * SC_NSCAL + SC_DONE] = 1.0;
}

// member-batched cycles: relative residual sqrt(dot) / nrm of every member
// (dot in its SC_T0); a member converges once at or below tol (sticky)
__global__ void mg_member_test_kernel(int M, const double* __restrict__ scal, double tol, double* __restrict__ stat,
                                      int init_nrm) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  double* done = stat;
  double* err = stat + M;
  double* nrm = stat + 2 * M;
  const double d = scal[(int64_t)k * SC_NSCAL + SC_T0];
  if (init_nrm) {  // nrm_f = l2(f), 1 if zero (lin_solve.f90:73)
    const double nf = sqrt(d);
    nrm[k] = nf == 0.0 ? 1.0 : nf;
    done[k] = 0.0;
    return;
  }
  err[k] = sqrt(d) / nrm[k];
  if (done[k] == 0.0 && err[k] <= tol) done[k] = 1.0;
}

// out := v on every cell of the array (work fields; not a memset, so the
// sequence stays capturable in a graph)
template <class F>
__global__ void mg_fill_kernel(int64_t n, double v, F* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (F)v;
}


// ---------------------------------------------------------------------------
// learned correction: features of 8x8 node patches and the scattered output
// ---------------------------------------------------------------------------
constexpr double LC_DEPTH_SCALE = 4000.0;  // m

// one wave per patch, one lane per node: row p of the features is
// [r / sigma | (|topo| + eta_old) / 4000 m | sin(latitude)], sigma the RMS of
// r over the patch (0 for a zero patch, then the first block is 0)
template <bool BF16>
__global__ void lc_features_kernel(int npatch, const int64_t* __restrict__ cells, const double* __restrict__ r,
                                   const double* __restrict__ topo, const double* __restrict__ eo,
                                   const double* __restrict__ fnode, double* __restrict__ sigma,
                                   float* __restrict__ outf, uint16_t* __restrict__ outb, int64_t ms) {
  const int p = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (p >= npatch) return;
  // member blockIdx.y: its fields at member * ms, its rows after the previous member's
  r = mem(r, ms), topo = mem(topo, ms), eo = mem(eo, ms), sigma = mem(sigma, (int64_t)npatch);
  outf = mem(outf, (int64_t)npatch * 192), outb = mem(outb, (int64_t)npatch * 192);
  const int64_t c = cells[(int64_t)p * 64 + lane];
  const double v = r[c];
  double ss = v * v;
  for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
  const double sg = sqrt(ss / 64.0);
  if (lane == 0) sigma[p] = sg;
  const float f0 = sg > 0.0 ? (float)(v / sg) : 0.f;
  const float f1 = (float)((fabs(topo[c]) + eo[c]) / LC_DEPTH_SCALE), f2 = (float)fnode[c];
  const int64_t row = (int64_t)p * 192;
  if (BF16) {
    auto bf = [](float f) {
      uint32_t u = __float_as_uint(f);
      u += 0x7fffu + ((u >> 16) & 1u);
      return (uint16_t)(u >> 16);
    };
    outb[row + lane] = bf(f0);
    outb[row + 64 + lane] = bf(f1);
    outb[row + 128 + lane] = bf(f2);
  } else {
    outf[row + lane] = f0;
    outf[row + 64 + lane] = f1;
    outf[row + 128 + lane] = f2;
  }
}

// e(patch node) = sigma * y (other cells keep their value)
// e := the network's correction y sigma on the patch nodes, 0 on the cells
// outside every patch (the pole masters), over the node compute list so the
// halo copies are written in the same pass; slot[cell] = patch * 64 + node
__global__ void lc_scatter_kernel(int n, const int64_t* __restrict__ ncl, const int* __restrict__ slot,
                                  const float* __restrict__ y, const double* __restrict__ sigma,
                                  double* __restrict__ e, int npatch, int64_t ms) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  y = mem(y, (int64_t)npatch * 64), sigma = mem(sigma, (int64_t)npatch), e = mem(e, ms);
  const int s = slot[ncl[2 * t + 1]];
  e[ncl[2 * t]] = s >= 0 ? (double)y[s] * sigma[s / 64] : 0.0;
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
  const int64_t NC = c->info.ncells * c->nmembers;  // member-major for ensembles
  for (int f = 0; f < FS_NFIELDS; ++f) {
    if (int rc = wtg_dalloc<double>(c, &c->fs[f], NC)) return rc;
    WTG_HIPCHK(c, hipMemsetAsync(c->fs[f], 0, NC * sizeof(double), c->stream));
  }
  // solver scalars, dot partials and block counters: one slice per member
  const int M = c->nmembers;
  if (int rc = wtg_dalloc<double>(c, &c->d_part, (size_t)RED_BLOCKS * RED_MAX * M)) return rc;
  if (int rc = wtg_dalloc<unsigned>(c, &c->d_count, M)) return rc;
  WTG_HIPCHK(c, hipMemsetAsync(c->d_count, 0, M * sizeof(unsigned), c->stream));
  if (int rc = wtg_dalloc<double>(c, &c->d_scal, (size_t)SC_NSCAL * M)) return rc;
  WTG_HIPCHK(c, hipMemsetAsync(c->d_scal, 0, SC_NSCAL * M * sizeof(double), c->stream));
  WTG_HIPCHK(c, hipHostMalloc((void**)&c->h_scal, SC_NSCAL * M * sizeof(double), hipHostMallocDefault));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int fs_field(wtg_ctx* c, int f, int nuser = FS_NUSER) {
  if (f < 0 || f >= nuser) return wtg_fail(c, WTG_E_ARG, "free-surface field %d out of range [0,%d)", f, nuser);
  return 0;
}

// the solver entry points act on one member's fields
int fs_single(wtg_ctx* c) {
  if (c->nmembers != 1)
    return wtg_fail(c, WTG_E_ARG, "ensemble contexts support the free surface through wtg_rk4_split only");
  return 0;
}

int fs_need_params(wtg_ctx* c) {
  if (!c->fs_params_set) return wtg_fail(c, WTG_E_STATE, "wtg_fs_set_params must be called first");
  return 0;
}

double fs_coef(const wtg_ctx* c) { return c->fs_theta1 * c->fs_theta2 * (c->fs_dt * c->fs_dt); }

int launch_apply(wtg_ctx* c, const double* q, double* out, bool gate = false, double* out2 = nullptr) {
  hipLaunchKernelGGL(fs_apply_kernel, dim3(nblk(c->nncl), c->nmembers), dim3(FS_NT), 0, c->stream,
                     gate ? c->d_scal : nullptr, c->nncl, c->d_ncl, fs_geo(c), q, c->fs[FS_ETA_OLD], c->params.grav,
                     fs_coef(c), out, out2, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

template <class Fu, class Ff, class Fo>
int launch_residual(wtg_ctx* c, const Fu* u, const Ff* f, Fo* out) {
  hipLaunchKernelGGL((fs_residual_kernel<Fu, Ff, Fo>), dim3(nblk(c->nncl), c->nmembers), dim3(FS_NT), 0, c->stream,
                     c->mg_mdone, c->nncl, c->d_ncl, fs_geo(c), u, f, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c),
                     out, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

template <class F, bool Zero = false>
int launch_jacobi(wtg_ctx* c, const F* u, const F* f, double w, F* out) {
  hipLaunchKernelGGL((fs_jacobi_kernel<F, Zero>), dim3(nblk(c->nncl), c->nmembers), dim3(FS_NT), 0, c->stream,
                     c->mg_mdone, c->nncl, c->d_ncl, fs_geo(c), u, f, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c),
                     c->fs_dt * c->fs_dt, c->fs_theta1 * c->fs_theta2, w, out, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

template <class F1, class F2, class Fo>
int launch_lc(wtg_ctx* c, double k1, int i1, const F1* s1, double k2, int i2, const F2* s2, Fo* out,
              bool gate = false) {
  hipLaunchKernelGGL((fs_lc_kernel<F1, F2, Fo>), dim3(nblk(c->nncl), c->nmembers), dim3(FS_NT), 0, c->stream,
                     gate ? c->d_scal : nullptr, c->mg_mdone, c->nncl, c->d_ncl, k1, i1, s1, k2, i2, s2, c->d_scal, out,
                     c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

template <class Fi, class Fo>
int launch_copy(wtg_ctx* c, const Fi* s, Fo* out) {
  hipLaunchKernelGGL((fs_copy_kernel<Fi, Fo>), dim3(nblk(c->nncl), c->nmembers), dim3(FS_NT), 0, c->stream, c->nncl,
                     c->d_ncl, s, out, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

// up to three dot products into scal[o0..]
int launch_dots_g(wtg_ctx* c, bool gate, int post_op, int nd, const double* a0, const double* b0, int o0,
                  const double* a1 = nullptr, const double* b1 = nullptr, int o1 = 0, const double* a2 = nullptr,
                  const double* b2 = nullptr, int o2 = 0) {
  const int nb = std::min(RED_BLOCKS, (int)nblk(c->nred));
  // one launch finishing in the last block (its device-scope fence writes the
  // L2 back): small single-member passes only -- with many members the
  // blocks' fences cost more than the second launch
  if (nb <= c->lastblock_max && c->nmembers == 1) {
    hipLaunchKernelGGL(fs_dot1_kernel, dim3(nb, c->nmembers), dim3(FS_NT), 0, c->stream, gate ? c->d_scal : nullptr,
                       gate ? 1 : 0, c->nred, c->d_rcell, c->d_rw, nd, a0, b0, a1 ? a1 : a0, b1 ? b1 : b0,
                       a2 ? a2 : a0, b2 ? b2 : b0, c->d_part, o0, o1, o2, c->d_scal, post_op, c->d_count,
                       c->info.ncells);
    WTG_HIPCHK(c, hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(fs_dot_kernel, dim3(nb, c->nmembers), dim3(FS_NT), 0, c->stream, gate ? c->d_scal : nullptr,
                     c->nred, c->d_rcell, c->d_rw, nd, a0, b0, a1 ? a1 : a0, b1 ? b1 : b0, a2 ? a2 : a0,
                     b2 ? b2 : b0, c->d_part, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(fs_dot_final_kernel, dim3(1, c->nmembers), dim3(FS_NT), 0, c->stream, gate ? 1 : 0, nb, nd,
                     c->d_part, o0, o1, o2, c->d_scal, post_op);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_dots(wtg_ctx* c, int nd, const double* a0, const double* b0, int o0, const double* a1 = nullptr,
                const double* b1 = nullptr, int o1 = 0, const double* a2 = nullptr, const double* b2 = nullptr,
                int o2 = 0) {
  return launch_dots_g(c, false, OP_NONE, nd, a0, b0, o0, a1, b1, o1, a2, b2, o2);
}

int launch_scalar(wtg_ctx* c, int op, bool gate = false, double arg = 0.0) {
  hipLaunchKernelGGL(fs_scalar_kernel, dim3(1, c->nmembers), dim3(64), 0, c->stream, op, gate ? 1 : 0, arg, c->d_scal);
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

// bicgstab (lin_solve.f90:361-431) on u with right-hand side f.  The
// iterations are queued in chunks without host synchronisation: every kernel
// of an iteration is gated by the device flag that the convergence test sets,
// so the iterations queued after convergence do nothing and the iterates and
// the iteration count are those of the sequential loop.  An iteration is five
// launches: dp(Ap, res0) -> alph; s = res - alph Ap with As = L s; dp(As, s),
// dp(As, As) -> omga; the update of u, res = s - omga As and its two dot
// products -> err, rho, b; p = res + b (p - omga Ap) with Ap = L p (the dot
// passes of large levels add a one-block final kernel).  u is updated on its
// masters in place and its halo copies refreshed after the loop.
constexpr int BICG_CHUNK = 8;

// zero_u: start from u = 0 (u is overwritten); sync_end: wait for the
// final ghost copy of u before returning (callers that stay on the stream
// need not)
int bicgstab(wtg_ctx* c, double* u, const double* f, double tol, int maxit, int* iters, double* err_out,
             bool zero_u = false, bool sync_end = true) {
  double** W = c->fs + FS_W0;
  double *res = W[0], *res0 = W[1], *p = W[2], *s = W[3], *Ap = W[4], *As = W[5], *p_alt = W[6];
  RC(launch_dots_g(c, false, OP_NRMF, 1, f, f, SC_T0));  // nrm_f = l2(f)
  const int M = c->nmembers;  // an ensemble solves every member's problem in the same launches
  const int64_t ms = c->info.ncells;
  hipLaunchKernelGGL(fs_bicg_init_kernel, dim3(nblk(c->nncl), M), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl,
                     fs_geo(c), u, f, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c), zero_u ? 1 : 0, res, res0, p,
                     s, ms);
  WTG_HIPCHK(c, hipGetLastError());
  RC(launch_apply(c, p, Ap, false, As));  // Ap = As = L p
  RC(launch_dots(c, 1, res0, res, SC_RHO));
  RC(launch_scalar(c, OP_LOOP_INIT, false, tol));
  if (c->mg_mdone) {
    hipLaunchKernelGGL(mg_mark_done_kernel, dim3((c->nmembers + 63) / 64), dim3(64), 0, c->stream, c->nmembers,
                       c->mg_mdone, c->d_scal);
    WTG_HIPCHK(c, hipGetLastError());
  }
  int queued = 0;
  bool done = false;
  // one chunk of gated iterations; the pointer swap of p returns to its
  // start after an even count, so a full chunk can be replayed from a graph
  double *Ap_alt = W[7];
  const int nbred = std::min(RED_BLOCKS, (int)nblk(c->nred));
  auto chunk = [&](int k) -> int {
    for (int i = 0; i < k; ++i) {
      RC(launch_dots_g(c, true, OP_ALPH, 1, Ap, res0, SC_T0));
      hipLaunchKernelGGL(fs_bicg_s_apply_kernel, dim3(nblk(c->nncl), M), dim3(FS_NT), 0, c->stream, c->d_scal,
                         c->nncl, c->d_ncl, fs_geo(c), res, Ap, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c),
                         c->d_scal, s, As, ms);  // s = res - alph Ap; As = L s
      WTG_HIPCHK(c, hipGetLastError());
      RC(launch_dots_g(c, true, OP_OMGA, 2, As, s, SC_T1, As, As, SC_T2));
      const bool last = nbred <= c->lastblock_max && M == 1;
      hipLaunchKernelGGL(fs_bicg_res_kernel, dim3(std::max((unsigned)nbred, nblk(c->nncl)), M), dim3(FS_NT), 0,
                         c->stream, c->d_scal, c->nmaster, c->nncl, c->d_ncl, c->nred, nbred, c->d_rcell, c->d_rw, u,
                         p, s, As, res0, c->d_scal, res, c->d_part, last ? c->d_count : nullptr,
                         ms);  // u, res, dots
      WTG_HIPCHK(c, hipGetLastError());
      if (!last) {  // err, the convergence test, rho and b
        hipLaunchKernelGGL(fs_dot_final_kernel, dim3(1, M), dim3(FS_NT), 0, c->stream, 1, nbred, 2, c->d_part,
                           SC_T0, SC_T3, 0, c->d_scal, OP_ERR_BETA);
        WTG_HIPCHK(c, hipGetLastError());
      }
      hipLaunchKernelGGL(fs_bicg_p_apply_kernel, dim3(nblk(c->nncl), M), dim3(FS_NT), 0, c->stream, c->d_scal,
                         c->nncl, c->d_ncl, fs_geo(c), res, p, Ap, c->fs[FS_ETA_OLD], c->params.grav, fs_coef(c),
                         c->d_scal, p_alt, Ap_alt, ms);  // p = res + b (p - omga Ap); Ap = L p
      WTG_HIPCHK(c, hipGetLastError());
      std::swap(p, p_alt);
      std::swap(Ap, Ap_alt);
    }
    return 0;
  };
  static_assert(BICG_CHUNK % 2 == 0, "a chunk must return p and Ap to their buffers");
  // the first chunk: the iterations the previous solve on this context took
  // to the same tolerance (repeated solves, such as the coarse level of
  // V-cycles or the free surface of successive steps, need about as many),
  // so few gated no-op iterations are queued past convergence
  const int first = c->bicg_hint > 0 && c->bicg_hint_tol == tol ? c->bicg_hint : BICG_CHUNK;
  while (queued < maxit && !done) {
    const int k = std::min(queued == 0 ? first : BICG_CHUNK, maxit - queued);
    if (c->fs_graph && k == BICG_CHUNK && M == 1) {
      // the operator's coefficients are kernel arguments of the graph
      const double key[2] = {fs_coef(c), c->params.grav};
      // the buffers that hold p and Ap at the chunk's start are part of the
      // key: after a chunk of odd length (a first chunk sized from the
      // previous solve) they are the other pair
      if (!c->bicg_graph || c->bicg_u != u || c->bicg_f != f || c->bicg_key[0] != key[0] ||
          c->bicg_key[1] != key[1] || c->bicg_p != p || c->bicg_Ap != Ap) {
        const double* p0 = p;
        const double* Ap0 = Ap;
        if (c->bicg_graph) {
          hipGraphExecDestroy(c->bicg_graph);
          c->bicg_graph = nullptr;
        }
        WTG_HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = chunk(k);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
        if (rc) {
          if (graph) hipGraphDestroy(graph);
          return rc;
        }
        if (ec != hipSuccess) return wtg_fail(c, WTG_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&c->bicg_graph, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
          c->bicg_graph = nullptr;
          return wtg_fail(c, WTG_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
        c->bicg_u = u;
        c->bicg_f = f;
        c->bicg_key[0] = key[0];
        c->bicg_key[1] = key[1];
        c->bicg_p = p0;
        c->bicg_Ap = Ap0;
      }
      WTG_HIPCHK(c, hipGraphLaunch(c->bicg_graph, c->stream));
    } else {
      RC(chunk(k));
    }
    queued += k;
    if (M == 1)
      WTG_HIPCHK(c, hipMemcpyAsync(c->h_scal + SC_ERR, c->d_scal + SC_ERR, (SC_TOL - SC_ERR) * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream));
    else
      WTG_HIPCHK(c, hipMemcpyAsync(c->h_scal, c->d_scal, (size_t)SC_NSCAL * M * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream));
    WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
    done = true;  // every member converged
    for (int m = 0; m < M; ++m) done = done && c->h_scal[m * SC_NSCAL + SC_DONE] != 0.0;
  }
  const int nghost = c->nncl - c->nmaster;
  if (nghost > 0) {
    hipLaunchKernelGGL(fs_ghost_copy_kernel, dim3(nblk(nghost), M), dim3(FS_NT), 0, c->stream, c->nmaster, c->nncl,
                       c->d_ncl, u, ms);
    WTG_HIPCHK(c, hipGetLastError());
  }
  if (sync_end) WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  // ensembles: the most iterations and the largest residual of the members
  int itmax = 0;
  double emax = 0.0;
  for (int m = 0; m < M; ++m) {
    itmax = std::max(itmax, (int)c->h_scal[m * SC_NSCAL + SC_ITER]);
    emax = std::max(emax, c->h_scal[m * SC_NSCAL + SC_ERR]);
  }
  *iters = done ? itmax : maxit + 1;
  *err_out = queued > 0 ? emax : 0.0;
  if (done) {
    c->bicg_hint = *iters;
    c->bicg_hint_tol = tol;
  }
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

int wtg_fs_ready(wtg_ctx* c) { return fs_ready(c); }

int wtg_fs_set_graph(wtg_ctx* c, int enable) {
  if (!c) return WTG_E_ARG;
  c->fs_graph = enable != 0;
  return 0;
}

// hierarchy of free-surface contexts, coarse (0) to fine (nl - 1)
struct wtg_mg {
  int nl = 0;
  std::vector<wtg_ctx*> lv;
  std::vector<double*> ovl;  // ovl[l], l >= 1: overlap areas between levels l and l - 1
  std::vector<wtg_grid_info> info;
  // learned correction on the finest level: 8x8 patches of owned nodes
  wtg_mlp* mlp = nullptr;
  int npatch = 0;
  int64_t* d_cells = nullptr;  // [npatch][64] cell of each patch node
  int* d_slot = nullptr;       // [ncells] patch * 64 + node of each patch cell, -1 elsewhere
  double* d_fnode = nullptr;   // sin(latitude) per cell
  double* d_sigma = nullptr;   // [npatch]
  float* d_feat = nullptr;     // [npatch][192] (feature extraction)
  // WTG_PRECISION_FP32_BULK: the V-cycles' correction (smoothing, transfers)
  // in FP32 fields; residuals of the solution and its update stay FP64
  int precision = WTG_PRECISION_FP64;
  // a hierarchy of ensemble contexts (nmembers M on every level): each
  // member's problem in the same launches, members gated once converged
  int M = 1;
  double* d_mstat = nullptr;  // [3][M]: converged flag, relative residual, l2 of f
  std::vector<int> m_cycles;  // per member, last cycles solve
  std::vector<double> m_resid;
};

namespace {

constexpr int MG_MAX_LEVELS = 16;

// errors of a level are reported through the finest level's context
int mg_err(wtg_mg* m, int rc, int l) {
  if (rc && l != m->nl - 1) m->lv[m->nl - 1]->err = m->lv[l]->err;
  return rc;
}

MgGeo mg_geo(const wtg_mg* m, int l) {
  const wtg_ctx *f = m->lv[l], *c = m->lv[l - 1];
  return MgGeo{m->ovl[l], f->d_hinv, c->d_hinv, f->info.ncells, f->info.S, f->info.n, c->info.S,
               f->info.south_pole, f->info.north_pole, c->info.south_pole, c->info.north_pole, c->info.ncells};
}

template <class F>
int mg_fill(wtg_ctx* c, F* a, double v) {
  const int64_t n = c->info.ncells * c->nmembers;
  hipLaunchKernelGGL(mg_fill_kernel<F>, dim3((unsigned)((n + FS_NT - 1) / FS_NT)), dim3(FS_NT), 0, c->stream, n, v, a);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

// wc := wavelets of s on level l, s_c := restriction of (s, wc) to level l - 1
template <class F>
int mg_forward(wtg_mg* m, int l, const F* s, F* wc, F* s_c) {
  wtg_ctx *f = m->lv[l], *c = m->lv[l - 1];
  const MgGeo g = mg_geo(m, l);
  hipLaunchKernelGGL(mg_wavelet_kernel<F>, dim3(nblk(f->nncl), f->nmembers), dim3(FS_NT), 0, f->stream, f->mg_mdone,
                     f->nncl, f->d_ncl, g, s, wc);
  WTG_HIPCHK(f, hipGetLastError());
  hipLaunchKernelGGL(mg_restrict_kernel<F>, dim3(nblk(c->nncl), f->nmembers), dim3(FS_NT), 0, f->stream, f->mg_mdone,
                     c->nncl, c->d_ncl, g, s, wc, s_c);
  WTG_HIPCHK(f, hipGetLastError());
  return 0;
}

// s_f := inverse transform of (s_c, wc) from level l - 1 to level l
// ZeroW: all wavelets zero (wc may be null)
template <class F, bool ZeroW = false>
int mg_inverse(wtg_mg* m, int l, const F* s_c, const F* wc, F* s_f) {
  wtg_ctx* f = m->lv[l];
  const MgGeo g = mg_geo(m, l);
  hipLaunchKernelGGL((mg_prolong_even_kernel<F, ZeroW>), dim3(nblk(f->nncl), f->nmembers), dim3(FS_NT), 0, f->stream,
                     f->mg_mdone, f->nncl, f->d_ncl, g, s_c, wc, s_f);
  WTG_HIPCHK(f, hipGetLastError());
  hipLaunchKernelGGL((mg_prolong_odd_kernel<F, ZeroW>), dim3(nblk(f->nncl), f->nmembers), dim3(FS_NT), 0, f->stream,
                     f->mg_mdone, f->nncl, f->d_ncl, g, wc, s_f);
  WTG_HIPCHK(f, hipGetLastError());
  return 0;
}

// a level's fields during a solve; every operation writes a fresh buffer
// (kernels over the compute list must not write what their halo threads
// read), so the current buffers are tracked as pointers
struct MgLv {
  wtg_ctx* c;
  double *u, *u_alt, *f, *res, *res_alt, *corr, *corr_alt;
};

template <class F>
int lv_jacobi(MgLv& v, F*& u, F*& alt, const F* f, double w, int n) {
  for (int i = 0; i < n; ++i) {
    RC(launch_jacobi(v.c, u, f, w, alt));
    std::swap(u, alt);
  }
  return 0;
}

// ||f - L u|| / nrm_f, read back
int lv_res_err(MgLv& v, const double* f, const double* u, double nrm_f, double* out) {
  double* r = v.c->fs[FS_W0];
  RC(launch_residual(v.c, u, f, r));
  RC(launch_dots(v.c, 1, r, r, SC_T0));
  double d;
  RC(read_scalar(v.c, SC_T0, &d));
  *out = std::sqrt(d) / nrm_f;
  return 0;
}

int lv_l2(MgLv& v, const double* a, double* out) {
  RC(launch_dots(v.c, 1, a, a, SC_T0));
  double d;
  RC(read_scalar(v.c, SC_T0, &d));
  *out = std::sqrt(d);
  return 0;
}

// zero-wavelet prolongation (prolong / prol_fun, lin_solve.f90:244-269) of
// level l - 1's coarse field into level l
template <class F>
int lv_prolong0(wtg_mg* m, int l, const F* coarse, F* fine) {
  return mg_inverse<F, true>(m, l, coarse, nullptr, fine);
}

// n Jacobi sweeps from u = 0 (the first reads no u; n = 0 leaves u = 0)
template <class F>
int lv_jacobi0(MgLv& v, F*& u, F*& alt, const F* f, double w, int n) {
  if (n == 0) return mg_fill(v.c, u, 0.0);
  RC((launch_jacobi<F, true>(v.c, u, f, w, alt)));
  std::swap(u, alt);
  return lv_jacobi(v, u, alt, f, w, n - 1);
}

std::vector<MgLv> mg_levels(wtg_mg* m) {
  std::vector<MgLv> V(m->nl);
  for (int l = 0; l < m->nl; ++l) {
    wtg_ctx* c = m->lv[l];
    V[l] = MgLv{c, c->fs[FS_ETA], c->fs[MG_U2], c->fs[FS_RHS], c->fs[MG_RES], c->fs[MG_RES2], c->fs[MG_CORR],
                c->fs[MG_CORR2]};
  }
  return V;
}

// results back into each level's WTG_FS_ETA
int mg_finish(std::vector<MgLv>& V) {
  for (auto& v : V)
    if (v.u != v.c->fs[FS_ETA]) RC(launch_copy(v.c, v.u, v.c->fs[FS_ETA]));
  WTG_HIPCHK(V.back().c, hipStreamSynchronize(V.back().c->stream));
  return 0;
}

int vcycle32(wtg_mg* m, std::vector<MgLv>& V, int l, const wtg_mg_params& P, double nrm_f, double* err,
             const double* r_in);

// v_cycle (lin_solve.f90:102-139) on level l: u := u + V-cycle correction,
// Jacobi post-smoothing, *err = ||f - L u|| / nrm_f
// r_in: f - L u of the current u when the caller holds it (the previous
// cycle's convergence test computed it), else null
int vcycle(wtg_mg* m, std::vector<MgLv>& V, int l, const wtg_mg_params& P, double nrm_f, double* err,
           const double* r_in = nullptr) {
  if (l == 0) r_in = nullptr;  // one level: the coarse solve reads V[0].res
  // one level has nothing to smooth: the FP64 coarse solve is the cycle
  if (m->precision == WTG_PRECISION_FP32_BULK && l >= 1) return vcycle32(m, V, l, P, nrm_f, err, r_in);
  if (!r_in) RC(launch_residual(V[l].c, V[l].u, V[l].f, V[l].res));
  for (int j = l; j >= 1; --j) {
    const double* rhs = j == l && r_in ? r_in : V[j].res;
    RC(lv_jacobi0(V[j], V[j].corr, V[j].corr_alt, rhs, 1.0, P.down_iter));  // from corr = 0
    RC(launch_residual(V[j].c, V[j].corr, rhs, V[j].res_alt));  // res := res - L corr
    std::swap(V[j].res, V[j].res_alt);
    RC(mg_forward(m, j, V[j].res, V[j].c->fs[MG_WC], V[j - 1].res));
  }
  int cit;
  double cerr;
  RC(bicgstab(V[0].c, V[0].corr, V[0].res, P.coarse_tol, P.coarse_iter, &cit, &cerr, true, false));
  for (int j = 1; j <= l; ++j) {
    double* pf = V[j].c->fs[MG_PF];
    RC(lv_prolong0(m, j, (const double*)V[j - 1].corr, pf));
    RC(launch_lc(V[j].c, 1.0, -1, V[j].corr, 1.0, -1, pf, V[j].corr_alt));
    std::swap(V[j].corr, V[j].corr_alt);
    RC(lv_jacobi(V[j], V[j].corr, V[j].corr_alt, V[j].res, 1.0, P.up_iter));
  }
  RC(launch_lc(V[l].c, 1.0, -1, V[l].u, 1.0, -1, V[l].corr, V[l].u_alt));
  std::swap(V[l].u, V[l].u_alt);
  RC(lv_jacobi(V[l], V[l].u, V[l].u_alt, V[l].f, 1.0, P.post_iter));
  if (m->M > 1) return 0;  // member-batched cycles test every member's residual in the caller
  return lv_res_err(V[l], V[l].f, V[l].u, nrm_f, err);
}

// FP32 fields of the correction cycle on every level of the hierarchy
int fs32_ready(wtg_mg* m) {
  for (wtg_ctx* c : m->lv) {
    if (c->fs32[0]) continue;
    const int64_t NC = c->info.ncells;
    for (int f = 0; f < M32_NFIELDS; ++f) {
      int rc = wtg_dalloc<float>(c, &c->fs32[f], NC);
      if (!rc && hipMemsetAsync(c->fs32[f], 0, NC * sizeof(float), c->stream) != hipSuccess)
        rc = wtg_fail(c, WTG_E_DEVICE, "hipMemsetAsync of an FP32 multigrid field failed");
      if (rc) {
        m->lv[m->nl - 1]->err = c->err;  // errors are reported through the finest level
        return rc;
      }
    }
  }
  return 0;
}

// The V-cycle of vcycle() as a defect correction with FP32 smoothing:
// R = f - L u in FP64, rounded to FP32; the correction equation L e = R is
// smoothed, restricted and prolonged in FP32 fields (FP64 arithmetic in
// registers), the coarsest level solved by FP64 BiCGSTAB; the post-smoothing
// sweeps on u are the same sweeps on e against R (Jacobi is affine:
// J(u + e; f) = u + J(e; f - L u)), and u := u + e in FP64, so the solution
// and its residual never round to FP32.
int vcycle32(wtg_mg* m, std::vector<MgLv>& V, int l, const wtg_mg_params& P, double nrm_f, double* err,
             const double* r_in) {
  std::vector<float *> corr(l + 1), corr_alt(l + 1), res(l + 1), res_alt(l + 1);
  for (int j = 0; j <= l; ++j) {
    float** b = V[j].c->fs32;
    corr[j] = b[M32_CORR], corr_alt[j] = b[M32_CORR2], res[j] = b[M32_RES], res_alt[j] = b[M32_RES2];
  }
  float* R = V[l].c->fs32[M32_R];
  if (r_in)
    RC(launch_copy(V[l].c, r_in, R));
  else
    RC(launch_residual(V[l].c, V[l].u, V[l].f, R));
  for (int j = l; j >= 1; --j) {
    const float* rhs = j == l ? R : res[j];
    RC(lv_jacobi0(V[j], corr[j], corr_alt[j], rhs, 1.0, P.down_iter));  // from corr = 0
    RC(launch_residual(V[j].c, corr[j], rhs, res_alt[j]));  // res := rhs - L corr
    std::swap(res[j], res_alt[j]);
    RC(mg_forward(m, j, (const float*)res[j], V[j].c->fs32[M32_WC], res[j - 1]));
  }
  int cit;
  double cerr;
  RC(launch_copy(V[0].c, (const float*)res[0], V[0].res));
  RC(bicgstab(V[0].c, V[0].corr, V[0].res, P.coarse_tol, P.coarse_iter, &cit, &cerr, true, false));
  RC(launch_copy(V[0].c, (const double*)V[0].corr, corr[0]));
  for (int j = 1; j <= l; ++j) {
    float* pf = V[j].c->fs32[M32_PF];
    RC(lv_prolong0(m, j, (const float*)corr[j - 1], pf));
    RC(launch_lc(V[j].c, 1.0, -1, (const float*)corr[j], 1.0, -1, (const float*)pf, corr_alt[j]));
    std::swap(corr[j], corr_alt[j]);
    RC(lv_jacobi(V[j], corr[j], corr_alt[j], (const float*)res[j], 1.0, P.up_iter));
  }
  RC(lv_jacobi(V[l], corr[l], corr_alt[l], (const float*)R, 1.0, P.post_iter));
  RC(launch_lc(V[l].c, 1.0, -1, (const double*)V[l].u, 1.0, -1, (const float*)corr[l], V[l].u_alt));
  std::swap(V[l].u, V[l].u_alt);
  return lv_res_err(V[l], V[l].f, V[l].u, nrm_f, err);
}

// FMG with V-cycles (lin_solve.f90:33-140)
int mg_fmg(wtg_mg* m, const wtg_mg_params& P, int* iters, double* res_out) {
  std::vector<MgLv> V = mg_levels(m);
  const int nl = m->nl;
  std::vector<double> nrm_f(nl), nrm1(nl);
  for (int l = 0; l < nl; ++l) {
    RC(lv_l2(V[l], V[l].f, &nrm_f[l]));
    if (nrm_f[l] == 0.0) nrm_f[l] = 1.0;
    RC(lv_res_err(V[l], V[l].f, V[l].u, nrm_f[l], &nrm1[l]));
    iters[l] = 0;
    res_out[l] = 0.0;
  }
  RC(bicgstab(V[0].c, V[0].u, V[0].f, P.coarse_tol, P.coarse_iter, &iters[0], &res_out[0]));
  for (int l = 1; l < nl; ++l) {
    if (nrm1[l] < P.fine_tol) continue;
    RC(lv_prolong0(m, l, (const double*)V[l - 1].u, V[l].u));
    RC(lv_jacobi(V[l], V[l].u, V[l].u_alt, V[l].f, 1.0, P.pre_iter));
    for (int it = 1; it <= P.max_vcycle; ++it) {
      RC(vcycle(m, V, l, P, nrm_f[l], &res_out[l]));
      iters[l] = it;
      if (res_out[l] <= P.fine_tol) break;
    }
  }
  return mg_finish(V);
}

// SRJ (lin_solve.f90:142-231)
int mg_srj(wtg_mg* m, const wtg_mg_params& P, int* iters, double* res_out) {
  constexpr int M = 8;
  const double k1 = 3e-2, k2 = 1.8, pi = std::acos(-1.0);
  std::vector<MgLv> V = mg_levels(m);
  const int nl = m->nl;
  double w[M];
  int max_it = P.max_srj_iter;
  if (max_it < M) {
    for (double& x : w) x = 1.0;
  } else {
    max_it = M * (max_it / M);
    for (int i = 1; i <= M; ++i)
      w[i - 1] = 2.0 / (k1 + k2 - (k2 - k1) * std::cos(pi * (2.0 * (double)i - 1.0) / (2.0 * (double)M)));
  }
  std::vector<double> nrm_f(nl);
  for (int l = 0; l < nl; ++l) {
    RC(lv_l2(V[l], V[l].f, &nrm_f[l]));
    if (nrm_f[l] == 0.0) nrm_f[l] = 1.0;
    iters[l] = 0;
    res_out[l] = 0.0;
  }
  RC(bicgstab(V[0].c, V[0].u, V[0].f, P.coarse_tol, P.coarse_iter, &iters[0], &res_out[0]));
  for (int l = 1; l < nl; ++l) {
    RC(lv_prolong0(m, l, (const double*)V[l - 1].u, V[l].u));
    int ii = 0;
    for (int it = 1; it <= max_it; ++it) {
      if (++ii > M) {
        if (res_out[l] < P.fine_tol) break;
        ii = 1;
      }
      RC(lv_jacobi(V[l], V[l].u, V[l].u_alt, V[l].f, w[ii - 1], 1));
      RC(lv_res_err(V[l], V[l].f, V[l].u, nrm_f[l], &res_out[l]));
      iters[l] = it;
      if (res_out[l] <= P.fine_tol) break;
    }
  }
  return mg_finish(V);
}

// features of field r on the finest level (BF16 into the MLP's input rows,
// or FP32 into d_feat)
int lc_features(wtg_mg* m, const double* r, bool to_mlp) {
  wtg_ctx* c = m->lv[m->nl - 1];
  uint16_t* x0 = nullptr;
  float* y = nullptr;
  int maxb = 0;
  if (to_mlp) wtg_mlp_buffers(m->mlp, &x0, &y, &maxb);
  const unsigned nb = (unsigned)((m->npatch + 3) / 4);
  // an ensemble hierarchy: the rows of member k follow member k - 1's
  const int64_t ms = c->info.ncells;
  if (to_mlp)
    hipLaunchKernelGGL(lc_features_kernel<true>, dim3(nb, c->nmembers), dim3(256), 0, c->stream, m->npatch, m->d_cells,
                       r, c->d_topo, c->fs[FS_ETA_OLD], m->d_fnode, m->d_sigma, (float*)nullptr, x0, ms);
  else
    hipLaunchKernelGGL(lc_features_kernel<false>, dim3(nb, c->nmembers), dim3(256), 0, c->stream, m->npatch,
                       m->d_cells, r, c->d_topo, c->fs[FS_ETA_OLD], m->d_fnode, m->d_sigma, m->d_feat,
                       (uint16_t*)nullptr, ms);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

// e := ghost-filled learned correction of residual r on the finest level
int lc_correction(wtg_mg* m, const double* r, double* e) {
  wtg_ctx* c = m->lv[m->nl - 1];
  uint16_t* x0;
  float* y;
  int maxb;
  wtg_mlp_buffers(m->mlp, &x0, &y, &maxb);
  const int rows = m->npatch * c->nmembers;  // every member's patches in one inference
  if (rows > maxb)
    return wtg_fail(c, WTG_E_ARG, "the network's batch (%d rows) is smaller than the %d patches of the members", maxb,
                    rows);
  RC(lc_features(m, r, true));
  if (int rc = wtg_mlp_infer_on(m->mlp, c->stream, rows))
    return wtg_fail(c, rc, "learned correction: %s", wtg_mlp_error_text(m->mlp));
  hipLaunchKernelGGL(lc_scatter_kernel, dim3(nblk(c->nncl), c->nmembers), dim3(FS_NT), 0, c->stream, c->nncl,
                     c->d_ncl, m->d_slot, y, m->d_sigma, e, m->npatch, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

// Cycles on the finest level from its current u: a plain V-cycle each, or
// (learned) first the network's correction e of the residual r, taken with
// the step a = <r, L e> / <L e, L e> that minimises ||r - a L e|| when a > 0
// (skipped otherwise, the V-cycle alone remains), then the V-cycle.
// iters[top] = cycles, res_out[top] the final relative residual; coarser
// levels serve the V-cycles and report 0.
int mg_cycles(wtg_mg* m, const wtg_mg_params& P, bool learned, int* iters, double* res_out) {
  std::vector<MgLv> V = mg_levels(m);
  const int l = m->nl - 1;
  for (int j = 0; j < m->nl; ++j) iters[j] = 0, res_out[j] = 0.0;
  double nrm_f;
  RC(lv_l2(V[l], V[l].f, &nrm_f));
  if (nrm_f == 0.0) nrm_f = 1.0;
  wtg_ctx* c = V[l].c;
  double err;
  RC(lv_res_err(V[l], V[l].f, V[l].u, nrm_f, &err));
  res_out[l] = err;
  // lv_res_err leaves f - L u of the current u in the top level's FS_W0
  for (int it = 1; it <= P.max_vcycle && err > P.fine_tol; ++it) {
    const double* r_top = c->fs[FS_W0];
    const bool step = learned && it <= P.learned_cycles;
    if (step) {
      double *r = c->fs[FS_W0], *e = c->fs[FS_W0 + 2], *w = c->fs[FS_W0 + 3];
      RC(lc_correction(m, r, e));
      RC(launch_apply(c, e, w));
      // the step is computed and applied on the device (a = 0 leaves u unchanged)
      RC(launch_dots_g(c, false, OP_MRSTEP, 2, r, w, SC_T0, w, w, SC_T1));
      RC(launch_lc(c, 1.0, -1, V[l].u, 1.0, SC_MR, e, V[l].u_alt));
      std::swap(V[l].u, V[l].u_alt);
    }
    RC(vcycle(m, V, l, P, nrm_f, &err, step ? nullptr : r_top));  // the learned step moved u
    iters[l] = it;
    res_out[l] = err;
  }
  return mg_finish(V);
}

// mg_cycles for a hierarchy of ensemble contexts: every member's V-cycles
// (and learned steps) in the same launches.  After each cycle every member's
// residual is tested on the device; a member at or below fine_tol is
// converged and from then on every kernel of the cycle leaves its fields as
// they are (its solution buffers are copied through), so each member's
// iterates and cycle count are those of its own solve.  iters[top] / res_out
// [top]: the most cycles and the largest residual over the members; per
// member in m_cycles / m_resid.
int mg_cycles_members(wtg_mg* m, const wtg_mg_params& P, bool learned, int* iters, double* res_out) {
  std::vector<MgLv> V = mg_levels(m);
  const int l = m->nl - 1, M = m->M;
  wtg_ctx* c = V[l].c;
  for (int j = 0; j < m->nl; ++j) iters[j] = 0, res_out[j] = 0.0;
  double* done = m->d_mstat;
  const unsigned tb = (unsigned)((M + 63) / 64);
  auto test = [&](int init) -> int {
    RC(launch_dots(c, 1, init ? V[l].f : c->fs[FS_W0], init ? V[l].f : c->fs[FS_W0], SC_T0));
    hipLaunchKernelGGL(mg_member_test_kernel, dim3(tb), dim3(64), 0, c->stream, M, c->d_scal, P.fine_tol,
                       m->d_mstat, init);
    WTG_HIPCHK(c, hipGetLastError());
    return 0;
  };
  std::vector<double> h(2 * (size_t)M);
  auto read = [&]() -> int {
    WTG_HIPCHK(c, hipMemcpyAsync(h.data(), m->d_mstat, 2 * M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  };
  RC(test(1));  // nrm_f per member, every member active
  for (auto* x : m->lv) x->mg_mdone = done;
  auto finish = [&](int rc) {
    for (auto* x : m->lv) x->mg_mdone = nullptr;
    return rc;
  };
  // r = f - L u of every member into FS_W0, its relative residual
  if (int rc = launch_residual(c, V[l].u, V[l].f, c->fs[FS_W0])) return finish(rc);
  if (int rc = test(0)) return finish(rc);
  if (int rc = read()) return finish(rc);
  m->m_cycles.assign(M, 0);
  m->m_resid.assign(h.begin() + M, h.end());
  auto all_done = [&]() {
    for (int k = 0; k < M; ++k)
      if (h[k] == 0.0) return false;
    return true;
  };
  for (int it = 1; it <= P.max_vcycle && !all_done(); ++it) {
    const double* r_top = c->fs[FS_W0];
    const bool step = learned && it <= P.learned_cycles;
    if (step) {
      double *r = c->fs[FS_W0], *e = c->fs[FS_W0 + 2], *w = c->fs[FS_W0 + 3];
      if (int rc = lc_correction(m, r, e)) return finish(rc);
      if (int rc = launch_apply(c, e, w)) return finish(rc);
      if (int rc = launch_dots_g(c, false, OP_MRSTEP, 2, r, w, SC_T0, w, w, SC_T1)) return finish(rc);
      if (int rc = launch_lc(c, 1.0, -1, V[l].u, 1.0, SC_MR, e, V[l].u_alt)) return finish(rc);
      std::swap(V[l].u, V[l].u_alt);
    }
    double unused;
    if (int rc = vcycle(m, V, l, P, 1.0, &unused, step ? nullptr : r_top)) return finish(rc);
    if (int rc = launch_residual(c, V[l].u, V[l].f, c->fs[FS_W0])) return finish(rc);
    if (int rc = test(0)) return finish(rc);
    const std::vector<double> before(h.begin(), h.begin() + M);
    if (int rc = read()) return finish(rc);
    for (int k = 0; k < M; ++k)
      if (before[k] == 0.0) {  // active during this cycle
        m->m_cycles[k] = it;
        m->m_resid[k] = h[M + k];
      }
  }
  for (int k = 0; k < M; ++k) {
    iters[l] = std::max(iters[l], m->m_cycles[k]);
    res_out[l] = std::max(res_out[l], m->m_resid[k]);
  }
  // the converged members' coarse BiCGSTAB flags stay set until its next start
  return finish(mg_finish(V));
}

// the free-surface solve of a split-explicit step on the hierarchy: eta^n
// restricted to every coarser level as its operator's eta_old (the
// reference's coarse levels hold restrictions of sol), then cycles on the
// finest level from eta^n
struct MgSplit {
  wtg_mg* m;
  wtg_mg_params P;
  bool learned;
};

int mg_split_solve(void* user, wtg_ctx* c, int* iters) {
  auto* s = (MgSplit*)user;
  wtg_mg* m = s->m;
  for (int l = m->nl - 1; l >= 1; --l)
    RC(mg_forward(m, l, m->lv[l]->fs[FS_ETA_OLD], m->lv[l]->fs[MG_WC], m->lv[l - 1]->fs[FS_ETA_OLD]));
  std::vector<int> it(m->nl);
  std::vector<double> res(m->nl);
  if (m->M > 1)
    RC(mg_cycles_members(m, s->P, s->learned, it.data(), res.data()));
  else
    RC(mg_cycles(m, s->P, s->learned, it.data(), res.data()));
  *iters = it[m->nl - 1];
  (void)c;
  return 0;
}

int mg_check_level(wtg_mg* m, int fine) {
  wtg_ctx* top = m->lv[m->nl - 1];
  if (fine < 1 || fine >= m->nl) return wtg_fail(top, WTG_E_ARG, "fine level %d out of range [1,%d)", fine, m->nl);
  return 0;
}

}  // namespace

int wtg_fs_bicgstab_all(wtg_ctx* c, double tol, int maxit, int* iters, double* err) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  return bicgstab(c, c->fs[FS_ETA], c->fs[FS_RHS], tol, maxit, iters, err);
}

int wtg_fs_flux_div_async(wtg_ctx* c, int slot) {
  RC(fs_ready(c));
  if (c->nmembers != 1) return wtg_fail(c, WTG_E_ARG, "the free-surface fields are per member: use a single-member context");
  if (slot < 0 || slot >= c->nslots) return wtg_fail(c, WTG_E_ARG, "slot %d out of range [0,%d)", slot, c->nslots);
  const auto& s = c->slots[slot];
  // edge fluxes into the first three solver work fields (no solve is running)
  double *F0 = c->fs[FS_W0], *F1 = c->fs[FS_W0 + 1], *F2 = c->fs[FS_W0 + 2];
  const int64_t ne = (int64_t)10 * (c->info.n + 2) * (c->info.n + 2);
  hipLaunchKernelGGL(fs_edge_flux_kernel, dim3((unsigned)((ne + FS_NT - 1) / FS_NT)), dim3(FS_NT), 0, c->stream,
                     c->info.n, fs_geo(c), c->params.nlayers, s.m, s.u, F0, F1, F2);
  WTG_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(fs_div_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c), F0,
                     F1, F2, c->fs[FS_DIVF]);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int wtg_fs_rhs_async(wtg_ctx* c) {
  RC(fs_ready(c));
  RC(fs_need_params(c));
  if (c->nmembers != 1) return wtg_fail(c, WTG_E_ARG, "the free-surface fields are per member: use a single-member context");
  hipLaunchKernelGGL(fs_rhs_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, c->fs[FS_ETA],
                     c->fs[FS_DIVF], c->fs[FS_DIVF_OLD], c->fs_dt, c->fs_theta2, c->params.ref_density,
                     c->fs[FS_RHS]);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int wtg_fs_flux_div_into(wtg_ctx* c, int slot, double* flux3, double* divf) {
  const auto& s = c->slots[slot];
  const int64_t NC = c->info.ncells;
  const int64_t ne = (int64_t)10 * (c->info.n + 2) * (c->info.n + 2);
  hipLaunchKernelGGL(fs_edge_flux_kernel, dim3((unsigned)((ne + FS_NT - 1) / FS_NT)), dim3(FS_NT), 0, c->stream,
                     c->info.n, fs_geo(c), c->params.nlayers, s.m, s.u, flux3, flux3 + NC, flux3 + 2 * NC);
  WTG_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(fs_div_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, c->stream, c->nncl, c->d_ncl, fs_geo(c), flux3,
                     flux3 + NC, flux3 + 2 * NC, divf);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

bool wtg_mg_same_levels(const wtg_mg* m, wtg_ctx* const* lv, int n) {
  if (!m || m->nl != n) return false;
  for (int l = 0; l < n; ++l)
    if (m->lv[l] != lv[l]) return false;
  return true;
}

int wtg_mg_eta_transform(wtg_mg* m) {
  for (int l = 1; l < m->nl; ++l) {
    wtg_ctx* f = m->lv[l];
    const MgGeo g = mg_geo(m, l);
    hipLaunchKernelGGL(mg_wavelet_kernel<double>, dim3(nblk(f->nncl), f->nmembers), dim3(FS_NT), 0, f->stream,
                       (const double*)nullptr, f->nncl, f->d_ncl, g, (const double*)f->fs[FS_ETA], f->fs[MG_WC]);
    WTG_HIPCHK(f, hipGetLastError());
  }
  for (int l = 1; l < m->nl; ++l)
    RC(mg_err(m, mg_inverse(m, l, (const double*)m->lv[l - 1]->fs[FS_ETA], (const double*)m->lv[l]->fs[MG_WC],
                            m->lv[l]->fs[FS_ETA]),
              l));
  return 0;
}

// ---------------------------------------------------------------------------
// dot products that allocate nothing (wtg_fs_dots_noalloc): the block
// partials live in a file-scope device array instead of the context's
// buffers.  Host code never passes `scratch` itself to a kernel: its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device
// when the first context is created there (wtg_noalloc_scratch_init).
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[RED_BLOCKS * RED_MAX];  // 12 KiB
constexpr int NOALLOC_MAX_DEVICES = 64;
static double* g_scratch[NOALLOC_MAX_DEVICES] = {};

int wtg_noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return WTG_E_DEVICE;
  if (g_scratch[device]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return WTG_E_DEVICE;
  g_scratch[device] = (double*)ptr;
  return 0;
}

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
  const int64_t NC = c->info.ncells, n = NC * c->nmembers;
  std::vector<double> t(host, host + n);
  for (int64_t m = 0; m < n; m += NC)
    for (int64_t cc = 0; cc < NC; ++cc) {
      const int64_t s = c->node_src[cc];
      if (s >= 0) t[m + cc] = host[m + s];
    }
  WTG_HIPCHK(c, hipMemcpyAsync(c->fs[field], t.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_fs_download(wtg_ctx* c, int field, double* host) {
  RC(fs_ready(c));
  RC(fs_field(c, field));
  if (!host) return wtg_fail(c, WTG_E_ARG, "null array");
  WTG_HIPCHK(c, hipMemcpyAsync(host, c->fs[field], c->info.ncells * c->nmembers * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_elliptic_apply(wtg_ctx* c, int in_field, int out_field) {
  RC(fs_ready(c));
  RC(fs_single(c));
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
  RC(fs_single(c));
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
  RC(fs_single(c));
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

int wtg_fs_dots_noalloc(wtg_ctx* c, int nd, const double* const* a, const double* const* b, double* out) {
  if (!c) return WTG_E_ARG;
  if (nd < 1 || nd > RED_MAX)
    return wtg_fail(c, WTG_E_ARG, "%d dot products asked; the scratch holds the partial sums of at most %d", nd,
                    RED_MAX);
  if (!a || !b || !out) return wtg_fail(c, WTG_E_ARG, "null argument");
  for (int k = 0; k < nd; ++k)
    if (!a[k] || !b[k]) return wtg_fail(c, WTG_E_ARG, "null field of dot product %d", k);
  if (c->nmembers != 1) return wtg_fail(c, WTG_E_ARG, "ensemble contexts are not supported");
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES || !g_scratch[c->device])
    return wtg_fail(c, WTG_E_STATE, "no scratch address for device %d", c->device);
  if (c->nred < 1) return wtg_fail(c, WTG_E_STATE, "the level has no reduction cells");
  double* part = g_scratch[c->device];
  const int nb = std::min(RED_BLOCKS, (int)nblk(c->nred));
  // every partial that the final pass reads is written by the first pass:
  // nothing in the scratch has to start at zero
  hipLaunchKernelGGL(fs_dot_kernel, dim3(nb, 1), dim3(FS_NT), 0, c->stream, (const double*)nullptr, c->nred,
                     c->d_rcell, c->d_rw, nd, a[0], b[0], nd > 1 ? a[1] : a[0], nd > 1 ? b[1] : b[0],
                     nd > 2 ? a[2] : a[0], nd > 2 ? b[2] : b[0], part, c->info.ncells);
  WTG_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(fs_dot_final_kernel, dim3(1, 1), dim3(FS_NT), 0, c->stream, 0, nb, nd, (const double*)part, 0, 1,
                     2, out, (int)OP_NONE);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int wtg_fs_dot(wtg_ctx* c, int a_field, int b_field, double* out) {
  RC(fs_ready(c));
  RC(fs_single(c));
  RC(fs_field(c, a_field));
  RC(fs_field(c, b_field));
  if (!out) return wtg_fail(c, WTG_E_ARG, "null output");
  RC(launch_dots(c, 1, c->fs[a_field], c->fs[b_field], SC_T0));
  return read_scalar(c, SC_T0, out);
}

int wtg_flux_divergence(wtg_ctx* c, int slot) {
  RC(wtg_fs_flux_div_async(c, slot));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_rhs_elliptic(wtg_ctx* c) {
  RC(wtg_fs_rhs_async(c));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_elliptic_solve(wtg_ctx* c, int method, double rel_tol, int max_iter, int* iters, double* rel_resid) {
  RC(fs_ready(c));
  RC(fs_single(c));
  RC(fs_need_params(c));
  if (!iters || !rel_resid) return wtg_fail(c, WTG_E_ARG, "null output");
  if (max_iter < 0) return wtg_fail(c, WTG_E_ARG, "max_iter < 0");
  if (method == WTG_SOLVE_BICGSTAB) return bicgstab(c, c->fs[FS_ETA], c->fs[FS_RHS], rel_tol, max_iter, iters, rel_resid);
  if (method == WTG_SOLVE_JACOBI) return jacobi_solve(c, c->fs[FS_ETA], c->fs[FS_RHS], rel_tol, max_iter, iters, rel_resid);
  return wtg_fail(c, WTG_E_ARG, "unknown solver method %d", method);
}


void wtg_mg_default_params(wtg_mg_params* p) {
  if (!p) return;
  p->coarse_tol = 1e-3;
  p->fine_tol = 1e-3;
  p->coarse_iter = 50;
  p->max_vcycle = 3;
  p->pre_iter = 2;
  p->down_iter = 2;
  p->up_iter = 2;
  p->post_iter = 4;
  p->max_srj_iter = 200;
  p->learned_cycles = 1 << 30;
}

int wtg_mg_create(const wtg_grid* const* grids, wtg_ctx* const* levels, int nlevels, wtg_mg** out) {
  if (!grids || !levels || !out) return wtg_fail(nullptr, WTG_E_ARG, "null argument");
  *out = nullptr;
  if (nlevels < 1 || nlevels > MG_MAX_LEVELS)
    return wtg_fail(nullptr, WTG_E_ARG, "nlevels %d out of range [1,%d]", nlevels, MG_MAX_LEVELS);
  for (int l = 0; l < nlevels; ++l) {
    if (!grids[l] || !levels[l]) return wtg_fail(nullptr, WTG_E_ARG, "level %d: null grid or context", l);
    wtg_ctx* c = levels[l];
    if (c->info.n != grids[l]->lv.n) return wtg_fail(nullptr, WTG_E_ARG, "level %d: context and grid differ", l);
    if (l && grids[l]->lv.level != grids[l - 1]->lv.level + 1)
      return wtg_fail(nullptr, WTG_E_ARG, "levels %d and %d are not consecutive", l - 1, l);
    if (c->device != levels[0]->device) return wtg_fail(nullptr, WTG_E_ARG, "level %d is on another device", l);
    if (c->nmembers != levels[0]->nmembers)
      return wtg_fail(nullptr, WTG_E_ARG, "level %d has %d members, level 0 %d", l, c->nmembers, levels[0]->nmembers);
    if (!c->topo_set || !c->fs_params_set)
      return wtg_fail(nullptr, WTG_E_STATE, "level %d: topography and free-surface parameters must be set", l);
    for (int k = 0; k < l; ++k)
      if (levels[k] == c) return wtg_fail(nullptr, WTG_E_ARG, "level %d repeats a context", l);
  }
  auto m = std::make_unique<wtg_mg>();
  m->nl = nlevels;
  m->M = levels[0]->nmembers;
  m->lv.assign(levels, levels + nlevels);
  m->ovl.assign(nlevels, nullptr);
  wtg_ctx* top = levels[nlevels - 1];
  for (int l = 0; l < nlevels; ++l) {
    if (int rc = fs_ready(levels[l])) {
      g_create_error = levels[l]->err;
      return rc;
    }
    if (hipStreamSynchronize(levels[l]->stream) != hipSuccess)
      return wtg_fail(nullptr, WTG_E_DEVICE, "level %d: stream synchronisation failed", l);
  }
  for (int l = 1; l < nlevels; ++l) {
    std::vector<double> a;
    try {
      a = wtg::overlap_areas(grids[l]->lv, grids[l - 1]->lv);
    } catch (const std::exception& e) {
      wtg_mg_destroy(m.release());
      return wtg_fail(nullptr, WTG_E_ARG, "%s", e.what());
    }
    if (hipMalloc((void**)&m->ovl[l], a.size() * sizeof(double)) != hipSuccess) {
      m->ovl[l] = nullptr;
      wtg_mg_destroy(m.release());
      return wtg_fail(nullptr, WTG_E_DEVICE, "hipMalloc of %zu overlap areas failed", a.size());
    }
    if (hipMemcpy(m->ovl[l], a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      wtg_mg_destroy(m.release());
      return wtg_fail(nullptr, WTG_E_DEVICE, "overlap area upload failed");
    }
  }
  // 8x8 patches of the finest level's owned nodes (row-major within a patch)
  const wtg::Level& fl = grids[nlevels - 1]->lv;
  if (fl.n % 8 == 0) {
    const int np1 = fl.n / 8;
    m->npatch = 10 * np1 * np1;
    std::vector<int64_t> cells((size_t)m->npatch * 64);
    for (int lz = 0; lz < 10; ++lz)
      for (int pj = 0; pj < np1; ++pj)
        for (int pi = 0; pi < np1; ++pi) {
          const int64_t p = ((int64_t)lz * np1 + pj) * np1 + pi;
          for (int b = 0; b < 8; ++b)
            for (int a = 0; a < 8; ++a) cells[p * 64 + b * 8 + a] = (int64_t)fl.cell(lz, 8 * pi + a, 8 * pj + b);
        }
    const size_t NC = fl.ncells();
    std::vector<int> slot(NC, -1);
    for (size_t q = 0; q < cells.size(); ++q) slot[cells[q]] = (int)q;
    std::vector<double> fn(NC);
    for (size_t c = 0; c < NC; ++c) fn[c] = fl.X[2 * NC + c] / fl.radius;
    bool ok = hipMalloc((void**)&m->d_cells, cells.size() * sizeof(int64_t)) == hipSuccess &&
              hipMalloc((void**)&m->d_fnode, NC * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&m->d_sigma, (size_t)m->npatch * m->M * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&m->d_feat, (size_t)m->npatch * 192 * sizeof(float)) == hipSuccess &&
              hipMalloc((void**)&m->d_slot, NC * sizeof(int)) == hipSuccess;
    ok = ok && hipMemcpy(m->d_cells, cells.data(), cells.size() * sizeof(int64_t), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m->d_fnode, fn.data(), NC * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m->d_slot, slot.data(), NC * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
      wtg_mg_destroy(m.release());
      return wtg_fail(nullptr, WTG_E_DEVICE, "patch table upload failed");
    }
  }
  if (hipMalloc((void**)&m->d_mstat, 3 * (size_t)m->M * sizeof(double)) != hipSuccess) {
    m->d_mstat = nullptr;
    wtg_mg_destroy(m.release());
    return wtg_fail(nullptr, WTG_E_DEVICE, "hipMalloc of the member status failed");
  }
  for (auto* c : m->lv) c->stream = top->own_stream;
  *out = m.release();
  return 0;
}

void wtg_mg_destroy(wtg_mg* m) {
  if (!m) return;
  if (!m->lv.empty()) hipStreamSynchronize(m->lv.back()->stream);
  for (auto* c : m->lv) c->stream = c->own_stream;
  for (double* a : m->ovl)
    if (a) hipFree(a);
  for (void* a : {(void*)m->d_cells, (void*)m->d_fnode, (void*)m->d_sigma, (void*)m->d_feat, (void*)m->d_slot,
                  (void*)m->d_mstat})
    if (a) hipFree(a);
  delete m;
}

static int mg_check_params(wtg_ctx* top, const wtg_mg_params* p) {
  if (p->coarse_iter < 0 || p->max_vcycle < 0 || p->pre_iter < 0 || p->down_iter < 0 || p->up_iter < 0 ||
      p->post_iter < 0 || p->max_srj_iter < 0 || p->learned_cycles < 0)
    return wtg_fail(top, WTG_E_ARG, "negative iteration count");
  return 0;
}

int wtg_mg_solve(wtg_mg* m, int method, const wtg_mg_params* p, int* iters, double* rel_resid) {
  if (!m) return WTG_E_ARG;
  wtg_ctx* top = m->lv[m->nl - 1];
  if (!p || !iters || !rel_resid) return wtg_fail(top, WTG_E_ARG, "null argument");
  if (int rc = mg_check_params(top, p)) return rc;
  if (m->M > 1) {  // an ensemble hierarchy: every member's cycles batched
    if (method != WTG_SOLVE_VCYCLE && method != WTG_SOLVE_LEARNED)
      return wtg_fail(top, WTG_E_ARG, "an ensemble hierarchy solves with WTG_SOLVE_VCYCLE or WTG_SOLVE_LEARNED");
    if (method == WTG_SOLVE_LEARNED && !m->mlp)
      return wtg_fail(top, WTG_E_STATE, "no network attached (wtg_mg_set_learned)");
    return mg_cycles_members(m, *p, method == WTG_SOLVE_LEARNED, iters, rel_resid);
  }
  if (method == WTG_SOLVE_FMG) return mg_fmg(m, *p, iters, rel_resid);
  if (method == WTG_SOLVE_SRJ) return mg_srj(m, *p, iters, rel_resid);
  if (method == WTG_SOLVE_VCYCLE) return mg_cycles(m, *p, false, iters, rel_resid);
  if (method == WTG_SOLVE_LEARNED) {
    if (!m->mlp) return wtg_fail(top, WTG_E_STATE, "no network attached (wtg_mg_set_learned)");
    return mg_cycles(m, *p, true, iters, rel_resid);
  }
  return wtg_fail(top, WTG_E_ARG, "unknown multigrid method %d", method);
}

int wtg_mg_member_results(const wtg_mg* m, int* cycles, double* rel_resid) {
  if (!m || !cycles || !rel_resid) return WTG_E_ARG;
  if ((int)m->m_cycles.size() != m->M)
    return wtg_fail(m->lv[m->nl - 1], WTG_E_STATE, "no member-batched solve has run on this hierarchy");
  for (int k = 0; k < m->M; ++k) cycles[k] = m->m_cycles[k], rel_resid[k] = m->m_resid[k];
  return 0;
}

int wtg_mg_forward(wtg_mg* m, int fine, int s_field, int wc_field, int coarse_field) {
  if (!m) return WTG_E_ARG;
  RC(mg_check_level(m, fine));
  wtg_ctx *f = m->lv[fine], *c = m->lv[fine - 1];
  RC(mg_err(m, fs_field(f, s_field), fine));
  RC(mg_err(m, fs_field(f, wc_field), fine));
  RC(mg_err(m, fs_field(c, coarse_field), fine - 1));
  if (s_field == wc_field) return wtg_fail(m->lv[m->nl - 1], WTG_E_ARG, "wavelets must not overwrite the scalar");
  RC(mg_err(m, mg_forward(m, fine, f->fs[s_field], f->fs[wc_field], c->fs[coarse_field]), fine));
  WTG_HIPCHK(f, hipStreamSynchronize(f->stream));
  return 0;
}

int wtg_mg_inverse(wtg_mg* m, int fine, int coarse_field, int wc_field, int fine_field) {
  if (!m) return WTG_E_ARG;
  RC(mg_check_level(m, fine));
  wtg_ctx *f = m->lv[fine], *c = m->lv[fine - 1];
  RC(mg_err(m, fs_field(c, coarse_field), fine - 1));
  RC(mg_err(m, fs_field(f, wc_field), fine));
  RC(mg_err(m, fs_field(f, fine_field), fine));
  if (wc_field == fine_field) return wtg_fail(m->lv[m->nl - 1], WTG_E_ARG, "the scalar must not overwrite the wavelets");
  RC(mg_err(m, mg_inverse(m, fine, c->fs[coarse_field], f->fs[wc_field], f->fs[fine_field]), fine));
  WTG_HIPCHK(f, hipStreamSynchronize(f->stream));
  return 0;
}


int wtg_mg_restrict_ss(wtg_mg* m, int fine, int fine_field, int coarse_field) {
  if (!m) return WTG_E_ARG;
  if (int rc = mg_check_level(m, fine)) return rc;
  wtg_ctx *f = m->lv[fine], *c = m->lv[fine - 1];
  if (int rc = fs_field(f, fine_field)) return mg_err(m, rc, fine);
  if (int rc = fs_field(c, coarse_field)) return mg_err(m, rc, fine - 1);
  hipLaunchKernelGGL(mg_restrict_ss_kernel, dim3(nblk(c->nncl)), dim3(FS_NT), 0, f->stream, c->nncl, c->d_ncl,
                     mg_geo(m, fine), f->fs[fine_field], c->fs[coarse_field]);
  WTG_HIPCHK(f, hipGetLastError());
  WTG_HIPCHK(f, hipStreamSynchronize(f->stream));
  return 0;
}

int wtg_mg_npatches(const wtg_mg* m) { return m ? m->npatch : WTG_E_ARG; }

int wtg_mg_set_precision(wtg_mg* m, int precision) {
  if (!m) return WTG_E_ARG;
  wtg_ctx* top = m->lv[m->nl - 1];
  if (precision != WTG_PRECISION_FP64 && precision != WTG_PRECISION_FP32_BULK)
    return wtg_fail(top, WTG_E_ARG, "unknown precision mode %d", precision);
  if (precision == WTG_PRECISION_FP32_BULK && m->M > 1)
    return wtg_fail(top, WTG_E_ARG, "FP32 smoothing is not available for an ensemble hierarchy");
  if (precision == WTG_PRECISION_FP32_BULK)
    if (int rc = fs32_ready(m)) return rc;
  m->precision = precision;
  return 0;
}

int wtg_mg_set_learned(wtg_mg* m, wtg_mlp* mlp) {
  if (!m) return WTG_E_ARG;
  wtg_ctx* top = m->lv[m->nl - 1];
  if (mlp) {
    if (!m->npatch) return wtg_fail(top, WTG_E_ARG, "the finest level's n is not a multiple of 8");
    uint16_t* x0;
    float* y;
    int maxb;
    wtg_mlp_buffers(mlp, &x0, &y, &maxb);
    if (maxb < m->npatch * m->M)
      return wtg_fail(top, WTG_E_ARG, "the network takes %d rows, the finest level has %d patches (all members)",
                      maxb, m->npatch * m->M);
  }
  m->mlp = mlp;
  return 0;
}

int wtg_mg_features(wtg_mg* m, int field, float* feats, double* sigma, int64_t* cells) {
  if (!m) return WTG_E_ARG;
  wtg_ctx* top = m->lv[m->nl - 1];
  if (!m->npatch) return wtg_fail(top, WTG_E_ARG, "the finest level's n is not a multiple of 8");
  if (m->M > 1) return wtg_fail(top, WTG_E_ARG, "features of an ensemble hierarchy are not available");
  RC(fs_field(top, field));
  if (!feats || !sigma) return wtg_fail(top, WTG_E_ARG, "null output");
  RC(lc_features(m, top->fs[field], false));
  WTG_HIPCHK(top, hipMemcpyAsync(feats, m->d_feat, (size_t)m->npatch * 192 * sizeof(float), hipMemcpyDeviceToHost,
                                 top->stream));
  WTG_HIPCHK(top, hipMemcpyAsync(sigma, m->d_sigma, (size_t)m->npatch * sizeof(double), hipMemcpyDeviceToHost,
                                 top->stream));
  if (cells)
    WTG_HIPCHK(top, hipMemcpyAsync(cells, m->d_cells, (size_t)m->npatch * 64 * sizeof(int64_t),
                                   hipMemcpyDeviceToHost, top->stream));
  WTG_HIPCHK(top, hipStreamSynchronize(top->stream));
  return 0;
}


int wtg_mg_rk4_split(wtg_mg* m, int slot, int nsteps, int method, const wtg_mg_params* p, int* iters) {
  if (!m) return WTG_E_ARG;
  wtg_ctx* top = m->lv[m->nl - 1];
  if (!p || !iters) return wtg_fail(top, WTG_E_ARG, "null argument");
  if (int rc = mg_check_params(top, p)) return rc;
  if (method != WTG_SOLVE_VCYCLE && method != WTG_SOLVE_LEARNED)
    return wtg_fail(top, WTG_E_ARG, "split steps on a hierarchy take WTG_SOLVE_VCYCLE or WTG_SOLVE_LEARNED, not %d",
                    method);
  if (method == WTG_SOLVE_LEARNED && !m->mlp)
    return wtg_fail(top, WTG_E_STATE, "no network attached (wtg_mg_set_learned)");
  MgSplit s{m, *p, method == WTG_SOLVE_LEARNED};
  *iters = 0;
  return wtg_split_steps(top, slot, nsteps, mg_split_solve, &s, iters);
}

}  // extern "C"
