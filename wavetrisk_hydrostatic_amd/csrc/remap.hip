// Vertical remap of the Lagrangian layers (wtg_remap): remap_vertical_coordinates
// (remap.f90:18-75) of the incompressible layer model (remap_incompressible,
// remap.f90:147-216) with the reference's remap_type "0" (remap0,
// piecewise constant, remap.f90:259-285), "1" (remap1, piecewise linear
// with harmonic-mean slopes and ten minmod reconciliation sweeps,
// remap.f90:287-380), "2PPM" (remap2PPM, Colella-Woodward piecewise
// parabolic with limited slopes, remap.f90:382-495) or "PPR" (remapPPR, the
// reference's default: Engwirda's PPR library at order 2 with the mono
// limiter and loose end conditions, remap.f90:950-1010), "2S" (remap2S,
// parabolic spline, remap.f90:497-583), "2W" (remap2W, parabolic WENO,
// remap.f90:585-735) or "4" (remap4, parabolic WENO with power-law
// reconciliation, remap.f90:737-948).
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

// remapPPR indexes edges 1..N+1 in these arrays (hence RMAX + 2); d, r, r1
// serve remap2S, remap2W and remap4
struct Work {
  double FC[RMAX + 2], aL[RMAX + 2], aR[RMAX + 2], Hz[RMAX + 2], CF[RMAX + 2], FC1[RMAX + 2];
  double d[RMAX + 2], r[RMAX + 2], r1[RMAX + 2];
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
  // each sweep's interface values FC(k) come from the previous sweep's aR, aL;
  // one pass computes FC(k) just before cell k overwrites aR(k), with FC(k-1)
  // in a register (the same numbers as the reference's two loops)
  for (int iter = 1; iter <= 10; ++iter) {
    double fprev = aL[0];  // FC(0)
    for (int k = 1; k <= N; ++k) {
      const double fk = k < N ? 0.5 * (aR[k - 1] + aL[k]) : aR[N - 1];  // FC(k)
      const double v = vo[k - 1];
      const double dR = fk - v, dL = v - fprev;
      const double cff = dR * dL < 0.0 ? 0.0 : (fabs(dR) > fabs(dL) ? dL : dR);
      aR[k - 1] = v + cff;
      aL[k - 1] = v - cff;
      fprev = fk;
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

// remap2S, remap2W and remap4 index the reference's (1:N) and (0:N) arrays at
// the same positions; V(k) is var_old(k).  Their arrays are cleared first: a
// column with fewer layers than a scheme's stencil reads entries the reference
// leaves unset, and those do not reach its result.
#define V(k) vo[(k) - 1]

__device__ void clear_work(int N, Work& w) {
  for (int k = 0; k <= N + 1; ++k) w.FC[k] = w.aL[k] = w.aR[k] = w.Hz[k] = w.CF[k] = w.FC1[k] = w.d[k] = w.r[k] = w.r1[k] = 0.0;
}

// remap2S (remap.f90:497-583): parabolic spline with parabolic continuation
// at both ends (dL in CF, dR in FC1)
__device__ void remap2s(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  clear_work(N, w);
  double *Hz = w.Hz, *dL = w.CF, *dR = w.FC1, *FC = w.FC, *r = w.r;
  for (int k = 1; k <= N; ++k) Hz[k] = zo[k] - zo[k - 1];
  if (N >= 2) {
    double cff = Hz[1] / Hz[2];
    FC[1] = 1.0 + cff;
    r[0] = 2.0 * V(1) + cff * (V(1) + cff * V(2)) / FC[1];
    for (int k = 1; k <= N - 1; ++k) {
      cff = 1.0 / (2.0 * Hz[k] + Hz[k + 1] * (2.0 - FC[k]));
      FC[k + 1] = cff * Hz[k];
      r[k] = cff * (3.0 * (V(k + 1) * Hz[k] + V(k) * Hz[k + 1]) - Hz[k + 1] * r[k - 1]);
    }
    cff = Hz[N] / Hz[N - 1];
    const double cff1 = 1.0 + cff;
    r[N] = (cff * (V(N) + cff * V(N - 1)) + cff1 * (2.0 * V(N) - cff1 * r[N - 1])) / (cff1 * (1.0 - cff1 * FC[N]));
    for (int k = N - 1; k >= 0; --k) r[k] = r[k] - FC[k + 1] * r[k + 1];
  }
  for (int k = 1; k <= N; ++k) {
    dR[k] = 0.5 * (r[k] - r[k - 1]);
    dL[k] = 0.5 * (r[k] + r[k - 1]) - V(k);
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    const bool up = dz > 0.0;
    const double alpha = dz / (up ? Hz[k + 1] : Hz[k]);
    const double cff = up ? dR[k + 1] : dR[k], cff1 = up ? dL[k + 1] : -dL[k];
    FC[k] = dz * (r[k] + alpha * (cff - cff1 * (3.0 - 2.0 * fabs(alpha))));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k] * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
}

// parabolic WENO side limits of the inner cells shared by remap2W and remap4
// (remap.f90:605-639, 760-795); d is remap2W's FC or remap4's d
__device__ void weno_sides(int N, const double* vo, const double* Hz, const double* d, double* aL, double* aR,
                           double* dL, double* dR) {
  for (int k = 2; k <= N - 1; ++k) {
    double deltaR = Hz[k] * d[k], deltaL = Hz[k] * d[k - 1];
    if (deltaR * deltaL < 0.0) {
      deltaR = 0.0;
      deltaL = 0.0;
    }
    double cff = Hz[k - 1] + 2.0 * Hz[k] + Hz[k + 1];
    const double cffR = cff * d[k], cffL = cff * d[k - 1];
    if (fabs(deltaR) > fabs(cffL)) deltaR = cffL;
    if (fabs(deltaL) > fabs(cffR)) deltaL = cffR;
    cff = (deltaR - deltaL) / (Hz[k - 1] + Hz[k] + Hz[k + 1]);
    deltaR = deltaR - cff * Hz[k + 1];
    deltaL = deltaL + cff * Hz[k - 1];
    aR[k] = V(k) + deltaR;
    aL[k] = V(k) - deltaL;
    const double pr = 2.0 * deltaR - deltaL, pl = 2.0 * deltaL - deltaR;
    dR[k] = pr * pr;
    dL[k] = pl * pl;
  }
}

// remap2W (remap.f90:585-735): LIMIT_INTERIOR, no NEUMANN (dL in CF, dR in FC1)
__device__ void remap2w(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  clear_work(N, w);
  double *Hz = w.Hz, *aL = w.aL, *aR = w.aR, *dL = w.CF, *dR = w.FC1, *FC = w.FC, *r = w.r;
  constexpr double eps = 1e-8;
  for (int k = 1; k <= N; ++k) Hz[k] = zo[k] - zo[k - 1];
  for (int k = N - 1; k >= 1; --k) FC[k] = (V(k + 1) - V(k)) / (Hz[k + 1] + Hz[k]);
  weno_sides(N, vo, Hz, FC, aL, aR, dL, dR);
  aR[N] = V(N); aL[N] = V(N); dR[N] = 0.0; dL[N] = 0.0;
  aR[1] = V(1); aL[1] = V(1); dR[1] = 0.0; dL[1] = 0.0;
  for (int k = 1; k <= N - 1; ++k) {
    const double deltaL = fmax(dL[k], eps), deltaR = fmax(dR[k + 1], eps);
    r[k] = (deltaR * aR[k] + deltaL * aL[k + 1]) / (deltaR + deltaL);
  }
  r[N] = 2.0 * V(N) - r[N - 1];
  r[0] = 2.0 * V(1) - r[1];
  for (int k = 1; k <= N; ++k) {
    double deltaR = r[k] - V(k), deltaL = V(k) - r[k - 1];
    const double cffR = 2.0 * deltaR, cffL = 2.0 * deltaL;
    if (deltaR * deltaL < 0.0) {
      deltaR = 0.0;
      deltaL = 0.0;
    } else if (fabs(deltaR) > fabs(cffL)) {
      deltaR = cffL;
    } else if (fabs(deltaL) > fabs(cffR)) {
      deltaL = cffR;
    }
    aR[k] = V(k) + deltaR;
    aL[k] = V(k) - deltaL;
    dL[k] = 0.5 * (aR[k] - aL[k]);
    dR[k] = 0.5 * (aR[k] + aL[k]) - V(k);
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    const bool up = dz > 0.0;
    const double alpha = dz / (up ? Hz[k + 1] : -Hz[k]);
    const double cff = up ? aL[k + 1] : aR[k], cffL = up ? dL[k + 1] : -dL[k], cffR = up ? dR[k + 1] : dR[k];
    FC[k] = dz * (cff + alpha * (cffL - cffR * (3.0 - 2.0 * alpha)));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k] * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
}

// remap4 (remap.f90:737-948): parabolic WENO, then the power-law
// reconciliation of interface values and derivatives (quartic profiles), no
// NEUMANN (dL in CF, dR in FC1).  The Pade constants are single-precision
// literals in the reference (default real kind), widened to double.
__device__ void remap4(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  clear_work(N, w);
  double *Hz = w.Hz, *aL = w.aL, *aR = w.aR, *dL = w.CF, *dR = w.FC1, *d = w.d, *FC = w.FC, *r = w.r, *r1 = w.r1;
  constexpr double eps = 1e-8;
  const double c1 = (double)0.0763636363636363636f, c2 = (double)0.004329004329004329f;
  const double c3 = (double)0.0363636363636363636f;
  for (int k = 1; k <= N; ++k) Hz[k] = zo[k] - zo[k - 1];
  for (int k = N - 1; k >= 1; --k) {
    FC[k] = 1.0 / (Hz[k + 1] + Hz[k]);
    d[k] = FC[k] * (V(k + 1) - V(k));
  }
  weno_sides(N, vo, Hz, d, aL, aR, dL, dR);
  {
    aL[N] = aR[N - 1];
    aR[N] = 2.0 * V(N) - aL[N];
    const double pr = 2.0 * aR[N] + aL[N] - 3.0 * V(N), pl = 3.0 * V(N) - 2.0 * aL[N] - aR[N];
    dR[N] = pr * pr;
    dL[N] = pl * pl;
  }
  if (N >= 2) {
    aR[1] = aL[2];
    aL[1] = 2.0 * V(1) - aR[1];
    const double pr = 2.0 * aR[1] + aL[1] - 3.0 * V(1), pl = 3.0 * V(1) - 2.0 * aL[1] - aR[1];
    dR[1] = pr * pr;
    dL[1] = pl * pl;
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double deltaL = fmax(dL[k], eps), deltaR = fmax(dR[k + 1], eps);
    r1[k] = (deltaR * aR[k] + deltaL * aL[k + 1]) / (deltaR + deltaL);
  }
  r1[N] = 2.0 * V(N) - r1[N - 1];
  r1[0] = 2.0 * V(1) - r1[1];
  for (int k = 1; k <= N; ++k) {
    const double deltaR = r1[k] - V(k), deltaL = V(k) - r1[k - 1];
    double cff = deltaR * deltaL;
    cff = cff > eps ? (deltaR + deltaL) / cff : 0.0;
    double cffL = cff * deltaL, cffR = cff * deltaR;
    if (cffL > 3.0) {
      cffL = cffL * deltaL;
      cffR = 0.0;
    } else if (cffR > 3.0) {
      cffL = 0.0;
      cffR = cffR * deltaR;
    } else {
      cffL = 4.0 * deltaL - 2.0 * deltaR;
      cffR = 4.0 * deltaR - 2.0 * deltaL;
    }
    cff = 1.0 / Hz[k];
    dR[k] = cff * cffR;
    dL[k] = cff * cffL;
  }
  for (int k = N - 1; k >= 1; --k) {
    d[k] = FC[k] * (Hz[k + 1] * dL[k + 1] + Hz[k] * dR[k]);
    const double cffR = 8.0 * (dR[k] + 2.0 * dL[k]), cffL = 8.0 * (dL[k + 1] + 2.0 * dR[k + 1]);
    if (fabs(d[k]) > fabs(cffR)) d[k] = cffR;
    if (fabs(d[k]) > fabs(cffL)) d[k] = cffL;
    double Hdd, rr;
    if ((dL[k + 1] - dR[k]) * (V(k + 1) - V(k)) > 0.0) {
      Hdd = Hz[k] * (d[k] - dR[k]);
      rr = V(k) - r1[k - 1];
    } else {
      Hdd = Hz[k + 1] * (dL[k + 1] - d[k]);
      rr = r1[k + 1] - V(k + 1);
    }
    rr = fabs(rr);
    double Ampl = 0.2 * Hdd * rr;
    Hdd = fabs(Hdd);
    const double cff = rr * rr + c1 * Hdd * (rr + c2 * Hdd);
    Ampl = cff > eps ? Ampl * (rr + c3 * Hdd) / cff : 0.0;
    r[k] = r1[k] + Ampl;
  }
  r[0] = 2.0 * V(1) - r[1];
  r[N] = 2.0 * V(N) - r[N - 1];
  d[0] = d[1];
  d[N] = d[N - 1];
  for (int k = 1; k <= N; ++k) {
    aR[k] = r[k] + r[k - 1] - 2.0 * V(k);
    dR[k] = Hz[k] * (d[k] + d[k - 1]) - 2.0 * (r[k] - r[k - 1]);
    dL[k] = Hz[k] * (d[k] - d[k - 1]) - 6.0 * aR[k];
  }
  for (int k = 1; k <= N - 1; ++k) {
    const double dz = zn[k] - zo[k];
    const bool up = dz > 0.0;
    const double alpha = dz / (up ? Hz[k + 1] : -Hz[k]);
    const double cff = up ? aR[k + 1] : aR[k], cffR = up ? dR[k + 1] : -dR[k], cffL = up ? dL[k + 1] : dL[k];
    FC[k] = dz * (r[k] + 0.5 * dz * d[k] +
                  alpha * alpha * (cff - cffR * (0.5 - 0.25 * alpha) + cffL * (1.0 - alpha * (1.25 - 0.5 * alpha))));
  }
  FC[0] = 0.0;
  FC[N] = 0.0;
  for (int k = 1; k <= N; ++k) vn[k - 1] = (Hz[k] * V(k) + FC[k] - FC[k - 1]) / (zn[k] - zn[k - 1]);
}
#undef V

// ---- remapPPR (remap.f90:950-1010): Engwirda's PPR library at order 2 ----
// cubic edge values (ppr/p3e.f90; p1e.f90 for at most three cells) with loose
// one-sided quadratic end fits (pbc.f90 lbc/ubc), monotone parabolas in the
// cells (ppm.f90 ppm/ppmfn with plm.f90 plsv/plsc slopes), exact integration of
// the cell parabolas over the new cells and the compensated conservation fix
// (rmap1d.f90 rmap1d/imap1d).  The macros index as the library does: edges
// 1..N+1, cells 1..N.

// the reference's compiled x**k: products from the left
__device__ __forceinline__ double pw2(double x) { return x * x; }
__device__ __forceinline__ double pw3(double x) { return (x * x) * x; }
__device__ __forceinline__ double pw4(double x) { return ((x * x) * x) * x; }

// slv_3x3 (inv.f90:64-124, 400-445): adjugate over determinant, in place
__device__ bool solve3(const double A[3][3], double v[3], double eps) {
  const double a2233 = A[1][1] * A[2][2], a2332 = A[1][2] * A[2][1], a2133 = A[1][0] * A[2][2];
  const double a2331 = A[1][2] * A[2][0], a2132 = A[1][0] * A[2][1], a2231 = A[1][1] * A[2][0];
  const double det = A[0][0] * (a2233 - a2332) - A[0][1] * (a2133 - a2331) + A[0][2] * (a2132 - a2231);
  if (!(fabs(det) > eps)) return false;
  const double a1233 = A[0][1] * A[2][2], a1332 = A[0][2] * A[2][1], a1223 = A[0][1] * A[1][2];
  const double a1322 = A[0][2] * A[1][1], a1133 = A[0][0] * A[2][2], a1331 = A[0][2] * A[2][0];
  const double a1123 = A[0][0] * A[1][2], a1321 = A[0][2] * A[1][0], a1132 = A[0][0] * A[2][1];
  const double a1231 = A[0][1] * A[2][0], a1122 = A[0][0] * A[1][1], a1221 = A[0][1] * A[1][0];
  const double t0 = ((a2233 - a2332) * v[0] + -(a1233 - a1332) * v[1] + (a1223 - a1322) * v[2]) / det;
  const double t1 = (-(a2133 - a2331) * v[0] + (a1133 - a1331) * v[1] + -(a1123 - a1321) * v[2]) / det;
  const double t2 = ((a2132 - a2231) * v[0] + -(a1132 - a1231) * v[1] + (a1122 - a1221) * v[2]) / det;
  v[0] = t0; v[1] = t1; v[2] = t2;
  return true;
}

// slv_4x4 (inv.f90:447-600): block LDU through 2 x 2 adjugates (the second
// singularity test repeats the first, as in the library)
__device__ bool solve4(const double A[4][4], double v[4], double eps) {
  const double adet = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  if (!(fabs(adet) > eps)) return false;
  const double i00 = A[1][1], i01 = -A[0][1], i10 = -A[1][0], i11 = A[0][0];
  // L = A21 adj(A11), U = adj(A11) A12, S = A22 - L A12 / adet (mul_2x2 with
  // scal +1, +1 and -1/adet)
  double L00 = 0.0, L01 = 0.0, L10 = 0.0, L11 = 0.0, U00 = 0.0, U01 = 0.0, U10 = 0.0, U11 = 0.0;
  L00 = L00 + (A[2][0] * i00 + A[2][1] * i10);
  L10 = L10 + (A[3][0] * i00 + A[3][1] * i10);
  L01 = L01 + (A[2][0] * i01 + A[2][1] * i11);
  L11 = L11 + (A[3][0] * i01 + A[3][1] * i11);
  U00 = U00 + (i00 * A[0][2] + i01 * A[1][2]);
  U10 = U10 + (i10 * A[0][2] + i11 * A[1][2]);
  U01 = U01 + (i00 * A[0][3] + i01 * A[1][3]);
  U11 = U11 + (i10 * A[0][3] + i11 * A[1][3]);
  const double sc = -1.0 / adet;
  double S[2][2] = {{A[2][2], A[2][3]}, {A[3][2], A[3][3]}};
  const double p00 = L00 * A[0][2] + L01 * A[1][2], p10 = L10 * A[0][2] + L11 * A[1][2];
  const double p01 = L00 * A[0][3] + L01 * A[1][3], p11 = L10 * A[0][3] + L11 * A[1][3];
  if (sc == 1.0) {
    S[0][0] = S[0][0] + p00; S[1][0] = S[1][0] + p10; S[0][1] = S[0][1] + p01; S[1][1] = S[1][1] + p11;
  } else if (sc == -1.0) {
    S[0][0] = S[0][0] - p00; S[1][0] = S[1][0] - p10; S[0][1] = S[0][1] - p01; S[1][1] = S[1][1] - p11;
  } else {
    S[0][0] = S[0][0] + sc * p00; S[1][0] = S[1][0] + sc * p10;
    S[0][1] = S[0][1] + sc * p01; S[1][1] = S[1][1] + sc * p11;
  }
  const double sdet = S[0][0] * S[1][1] - S[0][1] * S[1][0];
  v[2] = v[2] - (L00 * v[0] + L01 * v[1]) / adet;
  v[3] = v[3] - (L10 * v[0] + L11 * v[1]) / adet;
  const double t0 = (i00 * v[0] + i01 * v[1]) / adet, t1 = (i10 * v[0] + i11 * v[1]) / adet;
  v[0] = t0; v[1] = t1;
  const double t2 = (S[1][1] * v[2] + -S[0][1] * v[3]) / sdet, t3 = (-S[1][0] * v[2] + S[0][0] * v[3]) / sdet;
  v[2] = t2; v[3] = t3;
  v[0] = v[0] - (U00 * v[2] + U01 * v[3]) / adet;
  v[1] = v[1] - (U10 * v[2] + U11 * v[3]) / adet;
  return true;
}

#define PX(i) zo[(i) - 1]
#define PXN(i) zn[(i) - 1]
#define PF(i) vo[(i) - 1]
#define PDX(i) w.Hz[i]
#define PE(i) w.FC[i]

// loose end fit: the quadratic with the means of cells c-1, c, c+1, valued
// at edges c-1..c+1 (low end) or c..c+2 (high end)
__device__ void end_fit(bool low, int c, bool uni, const double* vo, double dmin, Work& w) {
  double xhat, h0, h1, h2;
  if (!uni) {
    xhat = fmax(PDX(c), dmin) * 0.5;
    h0 = PDX(c - 1); h1 = PDX(c); h2 = PDX(c + 1);
  } else {
    xhat = fmax(PDX(1), dmin) * 0.5;
    h0 = h1 = h2 = PDX(1);
  }
  const double s[4] = {-(h0 + h1 * 0.5) / xhat, -1.0, 1.0, (h2 + h1 * 0.5) / xhat};
  double A[3][3], v[3];
  const double h[3] = {h0, h1, h2};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    A[r][0] = s[r + 1] / 1.0 - s[r] / 1.0;
    A[r][1] = pw2(s[r + 1]) / 2.0 - pw2(s[r]) / 2.0;
    A[r][2] = pw3(s[r + 1]) / 3.0 - pw3(s[r]) / 3.0;
    v[r] = h[r] * PF(c - 1 + r) / xhat;
  }
  const int e0 = low ? c - 1 : c, o = low ? 0 : 1;
  if (solve3(A, v, 1.0e-14 * dmin)) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double x = s[k + o];
      PE(e0 + k) = 1.0 * v[0] + (x * 1.0) * v[1] + (pw2(x) * 1.0) * v[2];
    }
  } else if (low) {
    PE(c - 1) = PF(c - 1) * 1.0;
    PE(c) = PF(c - 1) * 0.5 + PF(c) * 0.5;
    PE(c + 1) = PF(c) * 0.5 + PF(c + 1) * 0.5;
  } else {
    PE(c) = PF(c - 1) * 0.5 + PF(c) * 0.5;
    PE(c + 1) = PF(c) * 0.5 + PF(c + 1) * 0.5;
    PE(c + 2) = PF(c + 1) * 1.0;
  }
}

__device__ __forceinline__ void parabola(double f0, double el, double er, double& c0, double& c1, double& c2) {
  c0 = 1.5 * f0 - 0.25 * (er + el);
  c1 = 0.5 * (er - el);
  c2 = -(1.5 * f0) + 0.75 * (er + el);
}

__device__ void remap_ppr(int N, double* vn, const double* zn, const double* vo, const double* zo, Work& w) {
  const int np = N + 1;
  // rmap1d: evenly spaced to 1e-14 keeps one spacing
  const double same = (PX(np) - PX(1)) / (np - 1), xtol = same * 1.0e-14;
  bool uni = true;
  for (int i = 1; i <= np - 1; ++i) {
    PDX(i) = PX(i + 1) - PX(i);
    if (fabs(PDX(i) - same) > xtol) uni = false;
  }
  // rcon1d: dmin = 1e-14 of the mean spacing
  double dmid = PDX(1);
  if (!uni) {
    for (int i = 2; i <= np - 1; ++i) dmid = dmid + PDX(i);
    dmid = dmid / (np - 1);
  }
  const double dmin = 1.0e-14 * dmid;
  // edge values
  if (N >= 2 && np <= 4) {
    for (int i = 2; i <= np - 1; ++i) {
      if (uni) {
        const double d = PDX(1) * 2.0;
        PE(i) = (PDX(1) * PF(i - 1) + PDX(1) * PF(i)) / d;
      } else {
        const double hl = fmax(PDX(i - 1), dmin), h0 = fmax(PDX(i), dmin);
        PE(i) = (h0 * PF(i - 1) + hl * PF(i)) / (hl + h0);
      }
    }
    PE(1) = PF(1);
    PE(np) = PF(np - 1);
  } else if (N >= 2) {
    end_fit(true, 2, uni, vo, dmin, w);
    end_fit(false, np - 2, uni, vo, dmin, w);
    // the four cells' spacings and values slide along in registers
    double h0 = PDX(1), h1 = PDX(2), h2 = PDX(3), f0 = PF(1), f1 = PF(2), f2 = PF(3);
    for (int i = 3; i <= np - 2; ++i) {
      const double h3 = PDX(i + 1), f3 = PF(i + 1);
      const double fw[4] = {f0, f1, f2, f3};
      h0 = h1; h1 = h2; h2 = h3;
      f0 = f1; f1 = f2; f2 = f3;
      if (uni) {
        PE(i) = (-(1.0 * fw[0]) + 7.0 * fw[1] + 7.0 * fw[2] - 1.0 * fw[3]) / 12.0;
        continue;
      }
      const double h[4] = {PDX(i - 2), h0, h1, h2};  // h0..h2 now hold PDX(i - 1 .. i + 1)
      const double xhat = 0.5 * fmax(h[1], dmin) + 0.5 * fmax(h[2], dmin);
      const double s[5] = {-(h[0] + h[1]) / xhat, -h[1] / xhat, 0.0, h[2] / xhat, (h[2] + h[3]) / xhat};
      double iv[5][4];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        iv[k][0] = s[k] / 1.0;
        iv[k][1] = pw2(s[k]) / 2.0;
        iv[k][2] = pw3(s[k]) / 3.0;
        iv[k][3] = pw4(s[k]) / 4.0;
      }
      double A[4][4], v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int p = 0; p < 4; ++p) A[r][p] = iv[r + 1][p] - iv[r][p];
        v[r] = h[r] * fw[r] / xhat;
      }
      PE(i) = solve4(A, v, 1.0e-14 * dmin) ? v[0] : fw[1] * 0.5 + fw[2] * 0.5;
    }
  }
  // monotone parabolas (coefficients of 1, s, s^2 on s in [-1, 1]) in aL, aR, CF
  // the neighbouring values, spacings and edges slide along in registers
  // (the end cells are their own outer neighbours)
  double fprev = PF(1), fcur = PF(1), hprev = PDX(1), hcur = hprev, ecur = PE(1);
  for (int i = 1; i <= N; ++i) {
    if (N == 1) {
      w.aL[1] = PF(1); w.aR[1] = 0.0; w.CF[1] = 0.0;
      break;
    }
    const double fnext = i < N ? PF(i + 1) : fcur, enext = PE(i + 1);
    const double hnext = uni ? hcur : (i < N ? PDX(i + 1) : hcur);
    const double f0 = fcur, fl = fprev, fr = fnext;
    const double hl_ = hprev, h0_ = hcur, hr_ = hnext;
    double el = ecur, er = enext;
    fprev = fcur; fcur = fnext;
    hprev = hcur; hcur = hnext;
    ecur = enext;
    // plsv / plsc: the limited centred slope
    double d0 = 0.0;
    {
      const double dm = f0 - fl, dp = fr - f0;
      if (dm * dp > 0.0) {
        double sl, sr;
        if (uni) {
          sl = (fl + f0) * 0.5;
          sr = (f0 + fr) * 0.5;
        } else {
          const double hl = hl_, h0 = h0_, hr = hr_;
          sl = (h0 * fl + hl * f0) / (hl + h0);
          sr = (hr * f0 + h0 * fr) / (h0 + hr);
        }
        d0 = 0.5 * (sr - sl);
        double scal = fmin(fabs(dm), fabs(dp)) / fmax(fabs(d0), 1.0e-14);
        scal = fmin(scal, 1.0);
        d0 = scal * d0;
      }
    }
    double u0, u1, u2, l0, l1, l2;
    parabola(f0, el, er, u0, u1, u2);
    if ((fr - f0) * (f0 - fl) < 0.0) {
      l0 = f0; l1 = 0.0; l2 = 0.0;
    } else {
      if ((fl - el) * (el - f0) <= 0.0) el = f0 - d0;
      if ((fr - er) * (er - f0) <= 0.0) er = f0 + d0;
      parabola(f0, el, er, l0, l1, l2);
      if (fabs(l2) > fabs(l1) * 0.5) {
        const double turn = -0.5 * l1 / l2;
        if (turn >= -1.0 && turn <= 0.0) {
          er = 3.0 * f0 - 2.0 * el;
          parabola(f0, el, er, l0, l1, l2);
        } else if (turn > 0.0 && turn <= 1.0) {
          el = 3.0 * f0 - 2.0 * er;
          parabola(f0, el, er, l0, l1, l2);
        }
      }
    }
    w.aL[i] = 0.0 * u0 + 1.0 * l0;
    w.aR[i] = 0.0 * u1 + 1.0 * l1;
    w.CF[i] = 0.0 * u2 + 1.0 * l2;
  }
  // imap1d: integrate the parabolas over the new cells
  int kmin = 1, kmax = 1, p0 = 1, p1 = 1;
  double* fnew = w.FC1;
  for (int k = 1; k <= N; ++k) {
    if (p1 < 1) p1 = 1;
    for (p0 = p1; p0 <= np - 1; ++p0)
      if (PX(p0 + 1) > PXN(k)) break;
    for (p1 = p0; p1 <= np - 1; ++p1)
      if (PX(p1) >= PXN(k + 1)) break;
    p1 = p1 - 1;
    const double khat = fmax(PXN(k + 1) - PXN(k), xtol);
    double acc = 0.0;
    for (int i = p0; i <= p1; ++i) {
      const double lo = fmax(PX(i), PXN(k)), hi = fmin(PX(i + 1), PXN(k + 1));
      const double mid = PX(i + 1) * 0.5 + PX(i) * 0.5, half = PX(i + 1) * 0.5 - PX(i) * 0.5;
      const double sl = (lo - mid) / fmax(half, xtol), sh = (hi - mid) / fmax(half, xtol);
      const double w0 = sh / 1.0 - sl / 1.0, w1 = pw2(sh) / 2.0 - pw2(sl) / 2.0, w2 = pw3(sh) / 3.0 - pw3(sl) / 3.0;
      double in = w0 * w.aL[i] + w1 * w.aR[i] + w2 * w.CF[i];
      in = in * half;
      acc = acc + in;
    }
    fnew[k] = acc / khat;
    if (fnew[k] > fnew[kmax])
      kmax = k;
    else if (fnew[k] < fnew[kmin])
      kmin = k;
  }
  // Kahan-Babuska-Neumaier sums of the old and new column contents; the
  // difference goes to the smallest (deficit) or largest (excess) new cell
  double so = 0.0, eo = 0.0, sn = 0.0, en = 0.0;
  for (int i = 1; i <= N; ++i) {
    const double b = (PX(i + 1) - PX(i)) * PF(i), t = so + b;
    eo = fabs(so) >= fabs(b) ? eo + ((so - t) + b) : eo + ((b - t) + so);
    so = t;
  }
  so = so + eo;
  for (int i = 1; i <= N; ++i) {
    const double b = (PXN(i + 1) - PXN(i)) * fnew[i], t = sn + b;
    en = fabs(sn) >= fabs(b) ? en + ((sn - t) + b) : en + ((b - t) + sn);
    sn = t;
  }
  sn = sn + en;
  const double err = so - sn;
  if (err > 0.0)
    fnew[kmin] = fnew[kmin] + err / (PXN(kmin + 1) - PXN(kmin));
  else if (err < 0.0)
    fnew[kmax] = fnew[kmax] + err / (PXN(kmax + 1) - PXN(kmax));
  for (int k = 1; k <= N; ++k) vn[k - 1] = fnew[k];
}
#undef PX
#undef PXN
#undef PF
#undef PDX
#undef PE

__device__ __forceinline__ void remap_col(int scheme, int N, double* vn, const double* zn, const double* vo,
                                          const double* zo, Work& w) {
  if (scheme == 0)
    remap0(N, vn, zn, vo, zo, w);
  else if (scheme == 1)
    remap1(N, vn, zn, vo, zo, w);
  else if (scheme == 2)
    remap2ppm(N, vn, zn, vo, zo, w);
  else if (scheme == 3)
    remap_ppr(N, vn, zn, vo, zo, w);
  else if (scheme == 4)
    remap2s(N, vn, zn, vo, zo, w);
  else if (scheme == 5)
    remap2w(N, vn, zn, vo, zo, w);
  else
    remap4(N, vn, zn, vo, zo, w);
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
  if (scheme < WTG_REMAP_PIECEWISE_CONSTANT || scheme > WTG_REMAP_QUARTIC_WENO)
    return wtg_fail(c, WTG_E_ARG,
                    "remap scheme %d is not built (0: remap0, 1: remap1, 2: remap2PPM, 3: remapPPR, 4: remap2S, "
                    "5: remap2W, 6: remap4)",
                    scheme);
  const int L = c->params.nlayers, M = c->nmembers;
  if (L > RMAX) return wtg_fail(c, WTG_E_ARG, "the remap takes at most %d layers, not %d", RMAX, L);
  // remap2S reads a second layer and remap4 a third one that the reference leaves unset below that
  if ((scheme == WTG_REMAP_PARABOLIC_SPLINE && L < 2) || (scheme == WTG_REMAP_QUARTIC_WENO && L < 3))
    return wtg_fail(c, WTG_E_ARG, "remap scheme %d needs at least %d layers, not %d", scheme,
                    scheme == WTG_REMAP_PARABOLIC_SPLINE ? 2 : 3, L);
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
