// Vertical remap of the Lagrangian layers (wtg_remap): remap_vertical_coordinates
// (remap.f90:18-75) of the incompressible layer model (remap_incompressible,
// remap.f90:147-216) with the reference's remap_type "0" (remap0,
// piecewise constant, remap.f90:259-285), "1" (remap1, piecewise linear
// with harmonic-mean slopes and ten minmod reconciliation sweeps,
// remap.f90:287-380) or "2PPM" (remap2PPM, Colella-Woodward piecewise
// parabolic with limited slopes, remap.f90:382-495).
//
// One thread per column (owned node or pole master; grid y = ensemble
// member): the old interface heights z_old(k) = z_s + sum of mass / rho0 and
// the target ones z_new(k) = a_vert(k) eta + b_vert(k) z_s, the buoyancy
// t / m remapped and the new mass rho0 dz, then each of the node's three
// edges: the column of interface heights averaged with the neighbour's, the
// velocity remapped (not at the poles).  Columns read the old masses from a
// copy taken before the launch, so no column sees another's update.  The
// column arrays are private (up to WTG_REMAP_MAX_LAYERS layers); FP64 in the
// reference's evaluation order without FMA contraction, so results equal the
// oracle bit for bit.  The kernel is HBM-bound on a single pass over the
// state (about 96 bytes read and written per node-layer).
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace {

constexpr int H = wtg::HALO;
constexpr int RMAX = WTG_REMAP_MAX_LAYERS;
constexpr int RT = 64;  // threads per block: the private column arrays are large

struct RemapArgs {
  int n, S, L, scheme;
  int64_t NC, spole, npole;
  const double *a, *b;  // [L + 1]
  const double* topo;   // [members][NC]
  const double* mold;   // [members][L][NC] masses before the remap
  double *m, *t, *u;    // the state slot
  double rho0;
};

struct Work {
  double FC[RMAX + 1], aL[RMAX + 1], aR[RMAX + 1], Hz[RMAX + 1], CF[RMAX + 1], FC1[RMAX + 1];
};

__device__ void remap0(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  double* FC = w.FC;
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    FC[k] = fmin(dz, 0.0) * vo[k - 1] + fmax(dz, 0.0) * vo[k];
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = ((zo[k] - zo[k - 1]) * vo[k - 1] + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
}

// the reference's arrays var(k), Hz(k), aL(k), aR(k) are [k - 1] here
__device__ void remap1(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  double *FC = w.FC, *aL = w.aL, *aR = w.aR, *Hz = w.Hz;
  for (int k = 0; k <= N; ++k) FC[k] = 0.0;  // one layer reads FC(1) before any store
  for (int k = 1; k <= N; ++k) Hz[k - 1] = zo[k] - zo[k - 1];
  for (int k = N - 1; k >= 1; --k) FC[k] = (vo[k] - vo[k - 1]) / (Hz[k] + Hz[k - 1]);
  FC[0] = FC[1];
  FC[N] = FC[N - 1];
  for (int k = 1; k <= N; ++k) {
    const double cff = FC[k] * FC[k - 1] < 0.0 ? 0.0 : 2 * FC[k] * FC[k - 1] / (FC[k] + FC[k - 1]);
    aR[k - 1] = vo[k - 1] + cff * Hz[k - 1];
    aL[k - 1] = vo[k - 1] - cff * Hz[k - 1];
  }
  for (int iter = 1; iter <= 10; ++iter) {
    for (int k = 1; k <= N - 1; ++k) FC[k] = 0.5 * (aR[k - 1] + aL[k]);
    FC[N] = aR[N - 1];
    FC[0] = aL[0];
    for (int k = 1; k <= N; ++k) {
      const double dR = FC[k] - vo[k - 1], dL = vo[k - 1] - FC[k - 1];
      const double cff = dR * dL < 0.0 ? 0.0 : (fabs(dR) > fabs(dL) ? dL : dR);
      aR[k - 1] = vo[k - 1] + cff;
      aL[k - 1] = vo[k - 1] - cff;
    }
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    const bool up = dz > 0.0;
    const double cff = up ? aL[k] : aR[k - 1];
    const double cff1 = up ? aR[k] - aL[k] : aR[k - 1] - aL[k - 1];
    const double dh = up ? Hz[k] : Hz[k - 1];
    FC[k] = dz * (cff + 0.5 * cff1 * dz / dh);
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k - 1] * vo[k - 1] + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
}

// remap2PPM: the reference's (0:N) arrays at [k], var(k) and Hz(k) at [k - 1];
// FC holds slopes, then curvatures, then the fluxes (in increasing k, as the
// reference overwrites it in place)
__device__ void remap2ppm(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  double *FC = w.FC, *aL = w.aL, *aR = w.aR, *Hz = w.Hz, *CF = w.CF, *FC1 = w.FC1;
  for (int k = 0; k <= N; ++k) FC[k] = CF[k] = FC1[k] = aL[k] = aR[k] = 0.0;
#define V(k) vo[(k) - 1]
#define HZ(k) Hz[(k) - 1]
  for (int k = 1; k <= N; ++k) HZ(k) = zo[k] - zo[k - 1];
  for (int k = 1; k <= N - 1; ++k) {
    const double cff = 1.0 / (HZ(k) + HZ(k + 1));
    CF[k] = cff * (V(k + 1) * HZ(k) + V(k) * HZ(k + 1));
    FC[k] = cff * (V(k + 1) - V(k));
  }
  for (int k = 2; k <= N - 1; ++k) {
    double cff = HZ(k) * ((2.0 * HZ(k - 1) + HZ(k)) * FC[k] + (2.0 * HZ(k + 1) + HZ(k)) * FC[k - 1]) /
                 (HZ(k - 1) + HZ(k) + HZ(k + 1));
    const double cffR = 2.0 * (V(k + 1) - V(k)), cffL = 2.0 * (V(k) - V(k - 1));
    if (cffR * cffL > 0.0) {
      if (fabs(cffL) < fabs(cff)) cff = cffL;
      if (fabs(cffR) < fabs(cff)) cff = cffR;
    } else {
      cff = 0.0;
    }
    FC1[k] = cff;
  }
  FC1[N] = HZ(N) * FC[N - 1];
  FC1[1] = HZ(1) * FC[1];
  for (int k = 1; k <= N - 1; ++k) {
    const int k1 = max(k - 1, 1), k2 = min(k + 2, N);
    const double cff = HZ(k1) + HZ(k) + HZ(k + 1) + HZ(k2);
    const double cffL = (HZ(k1) + HZ(k)) / (cff * (2.0 * HZ(k) + HZ(k + 1)));
    const double cffR = (HZ(k + 1) + HZ(k2)) / (cff * (HZ(k) + 2.0 * HZ(k + 1)));
    aR[k] = CF[k] + 2.0 * HZ(k) * HZ(k + 1) * (cffL - cffR) * FC[k] - cffL * HZ(k) * FC1[k + 1] +
            cffR * HZ(k + 1) * FC1[k];
    aL[k + 1] = aR[k];
  }
  aR[N] = 2.0 * V(N) - aL[N];
  aL[1] = 2.0 * V(1) - aR[1];
  for (int k = 1; k <= N; ++k) {
    CF[k] = 0.5 * (aR[k] - aL[k]);
    FC[k] = 0.5 * (aR[k] + aL[k]) - V(k);
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    const bool up = dz > 0.0;
    const double alpha = dz / (up ? HZ(k + 1) : -HZ(k));
    const double cff = up ? aL[k + 1] : aR[k];
    const double cffR = up ? CF[k + 1] : -CF[k];
    const double cffL = up ? FC[k + 1] : FC[k];
    FC[k] = dz * (cff + alpha * (cffR - cffL * (3.0 - 2.0 * alpha)));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (HZ(k) * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
#undef V
#undef HZ
}

__device__ __forceinline__ void remap_col(int scheme, int N, double* vn, const double* zn, const double* vo,
                                          const double* zo, Work& w) {
  if (scheme == 0)
    remap0(N, vn, zn, vo, zo, w);
  else if (scheme == 1)
    remap1(N, vn, zn, vo, zo, w);
  else
    remap2ppm(N, vn, zn, vo, zo, w);
}

// find_coordinates_incompressible (remap.f90:235-257) of column c
__device__ void coordinates(const RemapArgs& a, const double* topo, const double* mold, int64_t c, double* zn,
                            double* zo) {
  const double zs = topo[c];
  zo[0] = zs;
  for (int k = 1; k <= a.L; ++k) zo[k] = zo[k - 1] + (0.0 + mold[(int64_t)(k - 1) * a.NC + c]) / (a.rho0 * 1.0);
  const double eta = zo[a.L];
  for (int k = 0; k <= a.L; ++k) zn[k] = a.a[k] * eta + a.b[k] * zs;
}

__global__ __launch_bounds__(RT) void remap_kernel(RemapArgs a) {
  const int64_t nown = (int64_t)10 * a.n * a.n;
  const int64_t tid = (int64_t)blockIdx.x * RT + threadIdx.x;
  if (tid >= nown + 2) return;
  const int64_t NC = a.NC, s1 = (int64_t)blockIdx.y * a.L * NC;
  const double* topo = a.topo + (int64_t)blockIdx.y * NC;
  const double* mold = a.mold + s1;
  double *m = a.m + s1, *t = a.t + s1, *u = a.u + 3 * s1;
  int64_t c;
  bool pole = false;
  if (tid < nown) {
    const int lz = (int)(tid / ((int64_t)a.n * a.n)), r = (int)(tid % ((int64_t)a.n * a.n));
    c = (int64_t)lz * a.S * a.S + (int64_t)(r / a.n + H) * a.S + (r % a.n + H);
  } else {
    c = tid == nown ? a.spole : a.npole;
    pole = true;
  }
  double zo[RMAX + 1], zn[RMAX + 1], zeo[RMAX + 1], zen[RMAX + 1], vo[RMAX], vn[RMAX];
  Work w;
  const int L = a.L;
  coordinates(a, topo, mold, c, zn, zo);
  // buoyancy, then the new mass and mass-weighted buoyancy
  for (int k = 0; k < L; ++k) {
    const int64_t x = (int64_t)k * NC + c;
    vo[k] = (0.0 + t[x]) / (0.0 + m[x]);
  }
  remap_col(a.scheme, L, vn, zn, vo, zo, w);
  for (int k = 0; k < L; ++k) {
    const int64_t x = (int64_t)k * NC + c;
    const double rho_dz = (a.rho0 * 1.0) * (zn[k + 1] - zn[k]);
    m[x] = rho_dz - 0.0;
    t[x] = rho_dz * vn[k] - 0.0;
  }
  if (pole) return;  // velocities are not remapped at the poles
  const int64_t nb[3] = {c + 1, c + a.S + 1, c + a.S};
  for (int e = 0; e < 3; ++e) {
    coordinates(a, topo, mold, nb[e], zen, zeo);
    for (int k = 0; k <= L; ++k) {
      zen[k] = 0.5 * (zn[k] + zen[k]);
      zeo[k] = 0.5 * (zo[k] + zeo[k]);
    }
    for (int k = 0; k < L; ++k) vo[k] = u[((int64_t)k * 3 + e) * NC + c];
    remap_col(a.scheme, L, vn, zen, vo, zeo, w);
    for (int k = 0; k < L; ++k) u[((int64_t)k * 3 + e) * NC + c] = vn[k];
  }
}

}  // namespace

extern "C" int wtg_remap(wtg_ctx* c, int slot, int scheme, const double* a_vert, const double* b_vert) {
  if (!c || !a_vert || !b_vert) return WTG_E_ARG;
  if (slot < 0 || slot >= c->nslots) return wtg_fail(c, WTG_E_ARG, "slot %d out of range [0,%d)", slot, c->nslots);
  if (scheme != WTG_REMAP_PIECEWISE_CONSTANT && scheme != WTG_REMAP_PIECEWISE_LINEAR &&
      scheme != WTG_REMAP_PARABOLIC_PPM)
    return wtg_fail(c, WTG_E_ARG, "remap scheme %d is not built (0: remap0, 1: remap1, 2: remap2PPM)", scheme);
  const int L = c->params.nlayers, M = c->nmembers;
  if (L > RMAX) return wtg_fail(c, WTG_E_ARG, "the remap takes at most %d layers, not %d", RMAX, L);
  if (!c->topo_set) return wtg_fail(c, WTG_E_STATE, "wtg_set_topography must be called first");
  if (!c->d_remap_ab)
    if (int rc = wtg_dalloc<double>(c, &c->d_remap_ab, 2 * (RMAX + 1))) return rc;
  WTG_HIPCHK(c, hipMemcpyAsync(c->d_remap_ab, a_vert, (L + 1) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  WTG_HIPCHK(c, hipMemcpyAsync(c->d_remap_ab + RMAX + 1, b_vert, (L + 1) * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
  const auto& s = c->slots[slot];
  const int64_t NC = c->info.ncells;
  // update_bdry of the state, then the old masses into the first work slot
  if (int rc = wtg_ghost_fill_layers(c, L * M, s.m, s.t, s.u)) return rc;
  double* mold = c->slots[c->nslots].m;
  WTG_HIPCHK(c, hipMemcpyAsync(mold, s.m, (size_t)L * M * NC * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  RemapArgs a{c->info.n, c->info.S, L, scheme, NC, c->info.south_pole, c->info.north_pole, c->d_remap_ab,
              c->d_remap_ab + RMAX + 1, c->d_topo, mold, s.m, s.t, s.u, c->params.ref_density};
  const int64_t cols = (int64_t)10 * c->info.n * c->info.n + 2;
  hipLaunchKernelGGL(remap_kernel, dim3((unsigned)((cols + RT - 1) / RT), (unsigned)M), dim3(RT), 0, c->stream, a);
  WTG_HIPCHK(c, hipGetLastError());
  if (int rc = wtg_ghost_fill_layers(c, L * M, s.m, s.t, s.u)) return rc;
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
