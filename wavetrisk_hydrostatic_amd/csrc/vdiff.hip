// Vertical diffusion of an ocean column (wtg_vertical_diffusion):
// vertical_diffusion (vert_diffusion.f90:66-86) with the analytic eddy
// coefficients or the one-equation TKE closure (turbulent_diffusion,
// vert_diffusion.f90:88-264, with N_sq, l_scales, Richardson, Prandtl, Kv_tke
// and Kt_tke at 438-573), then the backward-Euler steps of the buoyancy
// (backwards_euler_temp, 266-355) and of the edge velocities
// (backwards_euler_velo, 357-424), each through LAPACK's dgtsv (dgtsv.f90).
//
// Two launches, one thread per column (grid y = ensemble member).  The first
// computes Kt, Kv and the TKE of every owned node and pole master; it reads
// the velocities of the node's six edges (its kinetic energy, io.f90:169-206),
// so it must finish before any velocity changes.  The second steps the
// buoyancy of the same columns and the velocities of each owned node's RT,
// DG and UP edges (with the node's Kv and the layer thicknesses averaged with
// the edge's other node).  Column arrays are private (up to
// WTG_REMAP_MAX_LAYERS layers); FP64 in the reference's evaluation order
// without FMA contraction.  The analytic profiles call exp, whose last bit may
// differ from the host library's; everything else equals the oracle bit for
// bit.  Both kernels are HBM-bound single passes over the column state.
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace {

constexpr int H = wtg::HALO;
constexpr int VMAX = WTG_REMAP_MAX_LAYERS;
constexpr int VT = 64;  // threads per block: the private column arrays are large

struct VdArgs {
  int n, S, L, istep;
  int64_t NC, spole, npole;
  double dt, rho0, grav;
  wtg_vdiff p;
  const double *len, *ped, *hinv, *topo, *eta;  // topo, eta [members][NC] (eta may be null)
  const double* force;                          // [members][6][NC]
  double *m, *t, *u;                            // the state slot
  double *kt, *kv, *tke;                        // [members][L + 1][NC], [members][L][NC]
};

// LAPACK dgtsv with one right-hand side on rows 1..n of 1-based arrays.  The
// current row's eliminated diagonal and right-hand side, and the last two
// solution values of the back substitution, ride in registers (the arrays
// live in scratch memory, and a reload of what the previous row just stored
// would put a memory round trip on every step of the sweep); every update is
// also stored, so the arrays hold what the library's would, an early return
// at a zero pivot included.
__device__ void dgtsv(int n, double* __restrict__ dl, double* __restrict__ d, double* __restrict__ du,
                      double* __restrict__ b) {
  double di = d[1], bi = b[1];
  for (int i = 1; i <= n - 1; ++i) {
    const double dli = dl[i], dui = du[i], dn = d[i + 1], bn = b[i + 1];
    if (fabs(di) >= fabs(dli)) {
      if (di == 0.0) return;
      const double fact = dli / di;
      di = dn - fact * dui;
      bi = bn - fact * bi;
      d[i + 1] = di;
      b[i + 1] = bi;
      if (i <= n - 2) dl[i] = 0.0;
    } else {  // interchange rows i and i + 1
      const double fact = di / dli;
      d[i] = dli;
      di = dui - fact * dn;
      d[i + 1] = di;
      if (i <= n - 2) {
        dl[i] = du[i + 1];
        du[i + 1] = -fact * dl[i];
      }
      du[i] = dn;
      b[i] = bn;
      bi = bi - fact * bn;
      b[i + 1] = bi;
    }
  }
  if (di == 0.0) return;  // d(n)
  double x1 = bi / di, x2 = 0.0;
  b[n] = x1;
  if (n > 1) {
    const double x = (b[n - 1] - du[n - 1] * x1) / d[n - 1];
    b[n - 1] = x;
    x2 = x1;
    x1 = x;
  }
  for (int i = n - 2; i >= 1; --i) {
    const double x = (b[i] - du[i] * x1 - dl[i] * x2) / d[i];
    b[i] = x;
    x2 = x1;
    x1 = x;
  }
}

struct SolveWork {
  double dg[VMAX + 2], dl[VMAX + 2], du[VMAX + 2], r[VMAX + 2];  // the rows, built only for the full dgtsv
  double Dp[VMAX + 2], Up[VMAX + 2], X[VMAX + 2];                // the elimination's diagonal, super-diagonal, rhs
};

// Solves the rows with dgtsv's arithmetic: while no row interchange or zero
// pivot occurs, the elimination keeps only its diagonal, super-diagonal and
// right-hand side (what dgtsv's no-interchange branch computes, the zeroed
// sub-diagonal entering the back substitution as 0.0 * x); otherwise the
// rows are built and dgtsv runs in full.  The solution is left in X[1..n].
// Rows gives reset(zs) and next(lo, dg, up, r), row after row from 1.
template <class Rows>
__device__ void solve_rows(int n, Rows& rows, double zs, SolveWork& w) {
  double lo, dg, up, r;
  rows.reset(zs);
  rows.next(lo, dg, up, r);
  double di = dg, xi = r, ui = up;
  w.Dp[1] = di;
  w.X[1] = xi;
  w.Up[1] = ui;
  bool fast = true;
  for (int i = 1; i <= n - 1; ++i) {
    rows.next(lo, dg, up, r);  // row i + 1: lo is dgtsv's dl(i)
    if (!(fabs(di) >= fabs(lo)) || di == 0.0) {
      fast = false;
      break;
    }
    const double fact = lo / di;
    di = dg - fact * ui;
    xi = r - fact * xi;
    ui = up;
    w.Dp[i + 1] = di;
    w.X[i + 1] = xi;
    w.Up[i + 1] = ui;
  }
  if (fast && di != 0.0) {
    double x1 = xi / di, x2 = 0.0;
    w.X[n] = x1;
    if (n > 1) {
      const double x = (w.X[n - 1] - w.Up[n - 1] * x1) / w.Dp[n - 1];
      w.X[n - 1] = x;
      x2 = x1;
      x1 = x;
    }
    for (int i = n - 2; i >= 1; --i) {
      const double x = (w.X[i] - w.Up[i] * x1 - 0.0 * x2) / w.Dp[i];
      w.X[i] = x;
      x2 = x1;
      x1 = x;
    }
    return;
  }
  rows.reset(zs);
  for (int k = 1; k <= n; ++k) {
    rows.next(lo, dg, up, r);
    if (k >= 2) w.dl[k - 1] = lo;
    w.dg[k] = dg;
    if (k <= n - 1) w.du[k] = up;
    w.r[k] = r;
  }
  dgtsv(n, w.dl, w.dg, w.du, w.r);
  for (int k = 1; k <= n; ++k) w.X[k] = w.r[k];
}

// the column of thread tid: owned nodes, then the south and north pole masters
__device__ __forceinline__ int64_t column(const VdArgs& a, int64_t tid, bool& pole) {
  const int64_t nown = (int64_t)10 * a.n * a.n;
  pole = tid >= nown;
  if (pole) return tid == nown ? a.spole : a.npole;
  const int lz = (int)(tid / ((int64_t)a.n * a.n)), r = (int)(tid % ((int64_t)a.n * a.n));
  return (int64_t)lz * a.S * a.S + (int64_t)(r / a.n + H) * a.S + (r % a.n + H);
}

__device__ __forceinline__ double prandtl(double Ri) {
  if (Ri < 0.2) return 1.0;
  if (Ri >= 0.2 && Ri <= 2.0) return 5.0 * Ri;
  return 10.0;
}

__device__ __forceinline__ double kv_tke(const wtg_vdiff& p, double e, double lk) {
  return fmin(p.Kv_max, fmax(p.C_k * lk * sqrt(e), p.Kv_min));
}

__device__ __forceinline__ double kt_tke(const wtg_vdiff& p, double Kv, double Nsq, double Ri) {
  double kt = fmin(p.Kt_max, fmax(Kv / prandtl(Ri), p.Kt_min));
  if (p.enhance_diff && Nsq <= p.Nsq_min) kt = p.Kt_enh;
  return kt;
}

__device__ __forceinline__ double free_surface(const VdArgs& a, const double* m, double zs, int64_t c) {
  double depth = 0.0;
  for (int k = 0; k < a.L; ++k) depth = depth + (0.0 + m[(int64_t)k * a.NC + c]) / (a.rho0 * 1.0);
  return depth + zs;
}

struct TkeWork {
  double dz[VMAX + 2], dU2[VMAX + 2], Nsq[VMAX + 2], e[VMAX + 2], leps[VMAX + 2], lk[VMAX + 2], dn[VMAX + 2];
};

// l_scales (vert_diffusion.f90:502-536) on [0..L]: the upward sweep of the
// mixing length from the bottom, then the downward one from the surface with
// the two length scales taken as it passes; the first-order length
// sqrt(2 e / max(N^2, Neps_sq)) is formed where each sweep needs it, and the
// running minima ride in registers (the same numbers as the reference's
// array passes)
__device__ void l_scales(const VdArgs& a, double tau, TkeWork& w) {
  const wtg_vdiff& p = a.p;
  const int L = a.L;
  double prev = p.l_0;
  w.dn[0] = prev;
  for (int l = 1; l <= L; ++l) {
    const double first = sqrt(2.0 * w.e[l] / fmax(w.Nsq[l], p.Neps_sq));
    w.dn[l] = prev = fmin(first, prev + w.dz[l]);
  }
  prev = fmax(p.l_0, p.kappa_VK * p.C_l / (a.rho0 * a.grav) * tau);  // l_up at the surface
  w.leps[L] = fmax(p.l_min, sqrt(prev * w.dn[L]));
  w.lk[L] = fmax(p.l_min, fmin(prev, w.dn[L]));
  for (int l = L - 1; l >= 0; --l) {
    const double first = sqrt(2.0 * w.e[l] / fmax(w.Nsq[l], p.Neps_sq));
    prev = fmin(first, prev + w.dz[l + 1]);
    w.leps[l] = fmax(p.l_min, sqrt(prev * w.dn[l]));
    w.lk[l] = fmax(p.l_min, fmin(prev, w.dn[l]));
  }
}

// Rows 1..L-1 of the TKE system (turbulent_diffusion), generated in order:
// the sources S1, S2 and the coefficients (coeff reads dzl at the row's own
// l) with the expressions of the reference's arrays
struct TkeRows {
  const VdArgs& a;
  const double *kv, *kt;
  int64_t c;
  const TkeWork& w;
  int l;
  __device__ void reset(double) { l = 0; }
  __device__ void next(double& lo, double& dg, double& up, double& r) {
    const wtg_vdiff& p = a.p;
    const int64_t NC = a.NC;
    const int last = a.L - 1;
    const double dt = a.dt;
    ++l;
    const double KV = kv[(int64_t)l * NC + c], KT = kt[(int64_t)l * NC + c];
    const double turb = w.e[l] == 0.0 ? 0.0 : p.C_eps * sqrt(w.e[l]) / w.leps[l];
    double S1 = KV * w.dU2[l] - KT * w.Nsq[l], S2;
    if (p.patankar && S1 <= 0.0) {
      S1 = KV * w.dU2[l];
      S2 = -turb - KT * w.Nsq[l] / w.e[l];
    } else {
      S2 = -turb;
    }
    const double dzl = 0.5 * (w.dz[l] + w.dz[l + 1]);
    up = l <= last - 1 ? -(dt * p.C_e * (0.5 * (KV + kv[(int64_t)(l + 1) * NC + c])) / (dzl * w.dz[l + 1])) : 0.0;
    lo = l >= 2 ? -(dt * p.C_e * (0.5 * (KV + kv[(int64_t)(l - 1) * NC + c])) / (dzl * w.dz[l])) : 0.0;
    if (l == 1)
      dg = 1.0 - up - dt * S2;
    else if (l < last)
      dg = 1.0 - (up + lo) - dt * S2;
    else
      dg = 1.0 - lo - dt * S2;
    r = w.e[l] + dt * S1;
  }
};

// turbulent_diffusion of every owned node and pole master
__global__ __launch_bounds__(VT) void vd_turbulent_kernel(VdArgs a) {
  const int64_t tid = (int64_t)blockIdx.x * VT + threadIdx.x;
  if (tid >= (int64_t)10 * a.n * a.n + 2) return;
  bool pole;
  const int64_t c = column(a, tid, pole);
  const int L = a.L;
  const int64_t NC = a.NC, mb = blockIdx.y;
  const double* m = a.m + mb * L * NC;
  const double* t = a.t + mb * L * NC;
  const double* u = a.u + mb * 3 * L * NC;
  double* kt = a.kt + mb * (L + 1) * NC;
  double* kv = a.kv + mb * (L + 1) * NC;
  double* tke = a.tke + mb * L * NC;
  const double zs = a.topo[mb * NC + c];
  const double tau = a.force[(mb * 6 + 2) * NC + c];
  const wtg_vdiff& p = a.p;
  if (!p.tke_closure) {
    const double eta = a.eta ? a.eta[mb * NC + c] : free_surface(a, m, zs, c);
    const double e20 = -20.0;
    double z = zs;
    kt[c] = p.Kt_min + p.Kt_max * exp(e20 * (z - eta) / p.mixed_layer);
    kv[c] = p.Kv_min + p.Kv_max * exp(e20 * (z - eta) / p.mixed_layer);
    for (int l = 1; l <= L; ++l) {
      z = z + (0.0 + m[(int64_t)(l - 1) * NC + c]) / (a.rho0 * 1.0);
      kt[(int64_t)l * NC + c] = p.Kt_min + p.Kt_max * exp(e20 * (z - eta) / p.mixed_layer);
      kv[(int64_t)l * NC + c] = p.Kv_min + p.Kv_max * exp(e20 * (z - eta) / p.mixed_layer);
    }
    return;
  }
  TkeWork w;
  // init_diffuse: dz and |u| (twice the TRiSK kinetic energy) per layer
  const int64_t W = c - 1, SW = c - a.S - 1, Sx = c - a.S;
  double umag = 0.0;  // |u| of the layer below, in a register
  for (int k = 1; k <= L; ++k) {
    w.dz[k] = (0.0 + m[(int64_t)(k - 1) * NC + c]) / (a.rho0 * 1.0);
    const double* uk = u + (int64_t)(k - 1) * 3 * NC;
    const double pRT = uk[c] * a.len[c], pDG = uk[NC + c] * a.len[NC + c], pUP = uk[2 * NC + c] * a.len[2 * NC + c];
    const double dRT = uk[c] * a.ped[c], dDG = uk[NC + c] * a.ped[NC + c], dUP = uk[2 * NC + c] * a.ped[2 * NC + c];
    const double pUPS = uk[2 * NC + Sx] * a.len[2 * NC + Sx], pDGSW = uk[NC + SW] * a.len[NC + SW];
    const double pRTW = uk[W] * a.len[W];
    const double dRTW = uk[W] * a.ped[W], dDGSW = uk[NC + SW] * a.ped[NC + SW], dUPS = uk[2 * NC + Sx] * a.ped[2 * NC + Sx];
    const double ke =
        (pUP * dUP + pDG * dDG + pRT * dRT + pUPS * dUPS + pDGSW * dDGSW + pRTW * dRTW) * a.hinv[c] / 4.0;
    const double um = sqrt(2.0 * ke);
    if (k >= 2) {  // dUdZ_sq at the interface below layer k
      const double q = (um - umag) / (0.5 * (w.dz[k - 1] + w.dz[k]));
      w.dU2[k - 1] = q * q;
    }
    umag = um;
  }
  w.dU2[0] = 0.0;
  w.dU2[L] = 0.0;
  for (int l = 0; l <= L; ++l) {
    const int k = l == 0 ? 1 : l == L ? L - 1 : l;  // N_sq at the interface above layer k
    const double bb = t[(int64_t)(k - 1) * NC + c] / (0.0 + m[(int64_t)(k - 1) * NC + c]);
    const double ba = t[(int64_t)k * NC + c] / (0.0 + m[(int64_t)k * NC + c]);
    w.Nsq[l] = a.grav * (ba - bb) / (0.5 * (w.dz[k] + w.dz[k + 1]));
  }
  w.e[0] = p.e_0;
  for (int l = 1; l <= L; ++l) w.e[l] = fmax(tke[(int64_t)(l - 1) * NC + c], p.e_min);
  l_scales(a, tau, w);
  if (a.istep == 1)
    for (int l = 0; l <= L; ++l) {
      const double Ri = w.Nsq[l] / (w.dU2[l] + p.eps_s);
      const double v = kv_tke(p, w.e[l], w.lk[l]);
      kv[(int64_t)l * NC + c] = v;
      kt[(int64_t)l * NC + c] = kt_tke(p, v, w.Nsq[l], Ri);
    }
  // the TKE system's backward-Euler step
  SolveWork sw;
  TkeRows rows{a, kv, kt, c, w, 0};
  solve_rows(L - 1, rows, 0.0, sw);
  w.e[0] = p.e_0;
  for (int l = 1; l <= L - 1; ++l) w.e[l] = p.patankar ? sw.X[l] : fmax(sw.X[l], p.e_min);
  w.e[L] = fmax(p.C_srf * tau / a.rho0, p.e_min_srf);
  // update_Kv_Kt
  l_scales(a, tau, w);
  for (int l = 0; l <= L; ++l) {
    const double Ri = w.Nsq[l] / (w.dU2[l] + p.eps_s);
    const double v = kv_tke(p, w.e[l], w.lk[l]);
    kv[(int64_t)l * NC + c] = v;
    kt[(int64_t)l * NC + c] = kt_tke(p, v, w.Nsq[l], Ri);
  }
  for (int l = 1; l <= L; ++l) tke[(int64_t)(l - 1) * NC + c] = fmax(w.e[l], p.e_min);
}

__device__ __forceinline__ double irradiance(const wtg_vdiff& p, double depth) {
  return p.Q_sr * (0.58 * exp(-depth / 0.35) + (1.0 - 0.58) * exp(-depth / 23.0));
}

// Rows of the backward-Euler systems, generated in order k = 1..L with the
// previous and next layer thicknesses carried in registers (each is the same
// expression as the reference's array entry, so the rows are identical).
// Buoyancy (backwards_euler_temp): e < 0; an edge's velocity: e = 0, 1, 2.
struct Rows {
  const VdArgs& a;
  const double *m, *t, *u, *K, *f;
  int64_t c, nb;
  int e;
  double eta, dt;
  int k;
  double dzp, dzc, dzn, z0, z1;  // dz(k-1), dz(k), dz(k+1); z(k-1), z(k)
  __device__ double dz(int kk) const {
    const int64_t x = (int64_t)(kk - 1) * a.NC;
    if (e < 0) return (0.0 + m[x + c]) / (a.rho0 * 1.0);
    return 0.5 * ((0.0 + m[x + c]) / (a.rho0 * 1.0) + (0.0 + m[x + nb]) / (a.rho0 * 1.0));
  }
  __device__ void reset(double zs) {
    k = 0;
    dzc = 0.0;
    dzn = dz(1);
    z1 = zs;
  }
  // row k + 1: sub-diagonal lo (k + 1 >= 2), diagonal dg, super-diagonal up (k + 1 <= L - 1), rhs r
  __device__ void next(double& lo, double& dg, double& up, double& r) {
    const int L = a.L;
    const wtg_vdiff& p = a.p;
    ++k;
    dzp = dzc;
    dzc = dzn;
    dzn = k < L ? dz(k + 1) : 0.0;
    z0 = z1;
    z1 = z0 + dzc;
    const double dzl_lo = 0.5 * (dzp + dzc), dzl_up = 0.5 * (dzc + dzn);
    lo = k >= 2 ? -(dt * K[(int64_t)(k - 1) * a.NC + c] / (dzl_lo * dzc)) : 0.0;
    up = k <= L - 1 ? -(dt * K[(int64_t)k * a.NC + c] / (dzl_up * dzc)) : 0.0;
    if (e < 0) {
      const int64_t x = (int64_t)(k - 1) * a.NC + c;
      const double b = t[x] / (0.0 + m[x]);
      if (k == 1) {
        dg = 1.0 - up;
        r = b + dt * (-f[c] + solar()) / dzc;
      } else if (k < L) {
        dg = 1.0 - (up + lo);
        r = b + dt * solar() / dzc;
      } else {
        // the reference's top row takes the row below's super-diagonal entry
        dg = 1.0 - (-(dt * K[(int64_t)(k - 1) * a.NC + c] / (dzl_lo * dzp)));
        r = b + dt * (f[a.NC + c] + p.Q_sr / (a.rho0 * p.c_p)) / dzc;
      }
    } else {
      r = u[((int64_t)(k - 1) * 3 + e) * a.NC + c];
      if (k == 1)
        dg = 1.0 - up + dt * p.bottom_friction / dzc;
      else if (k < L)
        dg = 1.0 - (up + lo);
      else
        dg = 1.0 - lo;
      if (k == L) r = r + dt * f[(3 + e) * a.NC + c] / dzc;
    }
  }
  __device__ double solar() const {
    const wtg_vdiff& p = a.p;
    return p.Q_sr * (irradiance(p, eta - z1) - irradiance(p, eta - z0)) / (a.rho0 * p.c_p);
  }
};

// backwards_euler_temp of every owned node and pole master, then
// backwards_euler_velo of the owned nodes' edges
__global__ __launch_bounds__(VT) void vd_solve_kernel(VdArgs a) {
  const int64_t tid = (int64_t)blockIdx.x * VT + threadIdx.x;
  if (tid >= (int64_t)10 * a.n * a.n + 2) return;
  bool pole;
  const int64_t c = column(a, tid, pole);
  const int L = a.L;
  const int64_t NC = a.NC, mb = blockIdx.y;
  const double* m = a.m + mb * L * NC;
  double* t = a.t + mb * L * NC;
  double* u = a.u + mb * 3 * L * NC;
  const double* f = a.force + mb * 6 * NC;
  SolveWork w;
  const double zs = a.topo[mb * NC + c];
  const double eta = a.eta ? a.eta[mb * NC + c] : free_surface(a, m, zs, c);
  Rows rows{a, m, t, u, a.kt + mb * (L + 1) * NC, f, c, c, -1, eta, a.dt};
  solve_rows(L, rows, zs, w);
  for (int k = 1; k <= L; ++k) {
    const int64_t x = (int64_t)(k - 1) * NC + c;
    t[x] = (0.0 + m[x]) * w.X[k];
  }
  if (pole) return;  // the poles carry no edges of their own
  const int64_t nb[3] = {c + 1, c + a.S + 1, c + a.S};
  for (int e = 0; e < 3; ++e) {
    Rows ru{a, m, t, u, a.kv + mb * (L + 1) * NC, f, c, nb[e], e, eta, a.dt};
    solve_rows(L, ru, zs, w);
    for (int k = 1; k <= L; ++k) u[((int64_t)(k - 1) * 3 + e) * NC + c] = w.X[k];
  }
}

}  // namespace

extern "C" void wtg_vdiff_defaults(wtg_vdiff* p) {
  if (!p) return;
  *p = wtg_vdiff{};
  p->C_e = 1.0;
  p->C_eps = 7.0e-1;
  p->C_l = 2.0e5;
  p->C_k = 1.0e-1;
  p->C_srf = 6.783e1;
  p->e_0 = 1.0e-6 / sqrt(2.0);
  p->e_min = 1.0e-6;
  p->e_min_srf = 1.0e-4;
  p->eps_s = 1.0e-20;
  p->kappa_VK = 4.0e-1;
  p->Kt_enh = 1.0;
  p->Kt_max = 1e-2;
  p->Kt_min = 1.2e-5;
  p->Kv_max = 1e-2;
  p->Kv_min = 1.2e-4;
  p->mixed_layer = -200.0;
  p->l_0 = 4.0e-2;
  p->l_min = 1.0e-2;
  p->Neps_sq = 1.0e-20;
  p->Nsq_min = 1.0e-12;
  p->Q_sr = 0.0;
  p->c_p = 1004.64;
}

extern "C" int wtg_set_vdiff(wtg_ctx* c, const wtg_vdiff* p) {
  if (!c || !p) return WTG_E_ARG;
  const int L = c->params.nlayers, M = c->nmembers;
  if (L > VMAX) return wtg_fail(c, WTG_E_ARG, "the vertical diffusion takes at most %d layers, not %d", VMAX, L);
  // backwards_euler_temp indexes a second layer and the TKE system a third one
  if (L < (p->tke_closure ? 3 : 2))
    return wtg_fail(c, WTG_E_ARG, "the vertical diffusion needs at least %d layers%s, not %d", p->tke_closure ? 3 : 2,
                    p->tke_closure ? " with the TKE closure" : "", L);
  const int64_t NC = c->info.ncells;
  if (!c->vd_kt) {
    if (int rc = wtg_dalloc<double>(c, &c->vd_kt, (size_t)M * (L + 1) * NC)) return rc;
    if (int rc = wtg_dalloc<double>(c, &c->vd_kv, (size_t)M * (L + 1) * NC)) return rc;
    if (int rc = wtg_dalloc<double>(c, &c->vd_tke, (size_t)M * L * NC)) return rc;
    if (int rc = wtg_dalloc<double>(c, &c->vd_force, (size_t)M * 6 * NC)) return rc;
    WTG_HIPCHK(c, hipMemsetAsync(c->vd_kt, 0, (size_t)M * (L + 1) * NC * sizeof(double), c->stream));
    WTG_HIPCHK(c, hipMemsetAsync(c->vd_kv, 0, (size_t)M * (L + 1) * NC * sizeof(double), c->stream));
    WTG_HIPCHK(c, hipMemsetAsync(c->vd_tke, 0, (size_t)M * L * NC * sizeof(double), c->stream));
    WTG_HIPCHK(c, hipMemsetAsync(c->vd_force, 0, (size_t)M * 6 * NC * sizeof(double), c->stream));
    WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->vd = *p;
  c->vd_set = true;
  return 0;
}

static int vd_ready(wtg_ctx* c) {
  if (!c) return WTG_E_ARG;
  if (!c->vd_set) return wtg_fail(c, WTG_E_STATE, "wtg_set_vdiff must be called first");
  return 0;
}

extern "C" int wtg_vdiff_set_coefficients(wtg_ctx* c, const double* Kt, const double* Kv, const double* tke) {
  if (int rc = vd_ready(c)) return rc;
  const int L = c->params.nlayers, M = c->nmembers;
  const size_t NC = (size_t)c->info.ncells, ni = (size_t)M * (L + 1) * NC, nl = (size_t)M * L * NC;
  if (Kt) WTG_HIPCHK(c, hipMemcpyAsync(c->vd_kt, Kt, ni * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (Kv) WTG_HIPCHK(c, hipMemcpyAsync(c->vd_kv, Kv, ni * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (tke) WTG_HIPCHK(c, hipMemcpyAsync(c->vd_tke, tke, nl * sizeof(double), hipMemcpyHostToDevice, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int wtg_vdiff_get_coefficients(wtg_ctx* c, double* Kt, double* Kv, double* tke) {
  if (int rc = vd_ready(c)) return rc;
  const int L = c->params.nlayers, M = c->nmembers;
  const size_t NC = (size_t)c->info.ncells, ni = (size_t)M * (L + 1) * NC, nl = (size_t)M * L * NC;
  if (Kt) WTG_HIPCHK(c, hipMemcpyAsync(Kt, c->vd_kt, ni * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (Kv) WTG_HIPCHK(c, hipMemcpyAsync(Kv, c->vd_kv, ni * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (tke) WTG_HIPCHK(c, hipMemcpyAsync(tke, c->vd_tke, nl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int wtg_vdiff_set_forcing(wtg_ctx* c, const double* bot, const double* top, const double* tau,
                                     const double* wind) {
  if (int rc = vd_ready(c)) return rc;
  const size_t NC = (size_t)c->info.ncells;
  const double* src[4] = {bot, top, tau, wind};
  const size_t off[4] = {0, 1, 2, 3}, cnt[4] = {1, 1, 1, 3};
  for (int mb = 0; mb < c->nmembers; ++mb)
    for (int q = 0; q < 4; ++q)
      if (src[q])
        WTG_HIPCHK(c, hipMemcpyAsync(c->vd_force + ((size_t)mb * 6 + off[q]) * NC, src[q] + (size_t)mb * cnt[q] * NC,
                                     cnt[q] * NC * sizeof(double), hipMemcpyHostToDevice, c->stream));
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int wtg_vertical_diffusion(wtg_ctx* c, int slot, double dt, int istep) {
  if (int rc = vd_ready(c)) return rc;
  if (slot < 0 || slot >= c->nslots) return wtg_fail(c, WTG_E_ARG, "slot %d out of range [0,%d)", slot, c->nslots);
  if (!c->topo_set) return wtg_fail(c, WTG_E_STATE, "wtg_set_topography must be called first");
  const double* eta = nullptr;
  if (c->vd.mode_split) {
    if (!c->fs[FS_ETA])
      return wtg_fail(c, WTG_E_STATE, "mode_split takes eta from the free surface, which has not been set");
    eta = c->fs[FS_ETA];
  }
  const int L = c->params.nlayers, M = c->nmembers;
  const auto& s = c->slots[slot];
  if (int rc = wtg_ghost_fill_layers(c, L * M, s.m, s.t, s.u)) return rc;
  VdArgs a{c->info.n, c->info.S, L, istep, c->info.ncells, c->info.south_pole, c->info.north_pole, dt,
           c->params.ref_density, c->params.grav, c->vd, c->d_len, c->d_ped, c->d_hinv, c->d_topo, eta, c->vd_force,
           s.m, s.t, s.u, c->vd_kt, c->vd_kv, c->vd_tke};
  const int64_t cols = (int64_t)10 * c->info.n * c->info.n + 2;
  const dim3 grid((unsigned)((cols + VT - 1) / VT), (unsigned)M);
  hipLaunchKernelGGL(vd_turbulent_kernel, grid, dim3(VT), 0, c->stream, a);
  WTG_HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(vd_solve_kernel, grid, dim3(VT), 0, c->stream, a);
  WTG_HIPCHK(c, hipGetLastError());
  if (int rc = wtg_ghost_fill_layers(c, L * M, s.m, s.t, s.u)) return rc;
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
