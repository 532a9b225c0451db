// MI355X kernels and C ABI of the TRiSK hot path (see include/wavetrisk_gpu.h).
//
// One RK stage is one launch of trisk_stage_kernel followed by one launch of
// ghost_fill_kernel.  The stage kernel evaluates trend_ml (multi_level.f90:5-47:
// cal_surf_press, integrate_pressure_up, step1 itype 0, post_step1,
// scalar_trend, du_source/Qperp, du_grad; ops.f90) on a 16 x 16 tile of owned
// nodes of one lozenge, for every layer bottom to top, and fuses the RK
// sub-step (time_integr.f90:206-225) into its stores:
//   out = q + h * trend(y)      (or out = trend(y) for wtg_trend)
// Per layer the block stages the tile's state plus a halo of one cell
// (low side) and two cells (high side) in LDS, computes potential vorticity
// on triangles, qe and mass fluxes on edges, kinetic energy / Bernoulli /
// Exner on nodes, and finally the trends of the owned nodes and edges.  The
// hydrostatic column scan (surface pressure, then pressure and geopotential
// carried upward through the layers) lives in registers across the layer
// loop, so a node's column is read once per stage.
//
// Arithmetic is FP64 in the reference's evaluation order and compiled
// without FMA contraction, so results match the CPU oracle bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "ctx.h"

namespace {

constexpr int H = wtg::HALO;  // 2
constexpr int T = 16;         // owned nodes per tile side
constexpr int R = T + 3;      // staged region: tile cells -1 .. T+1
constexpr int RR = R * R;     // 361
constexpr int NP = T + 1;     // pressure / Bernoulli cells per side: 0 .. T
constexpr int NT = T + 2;     // triangle / edge cells per side: -1 .. T
constexpr int THREADS = 256;

struct Geo {
  const double* len;    // [3][NC]
  const double* ped;    // [3][NC] local dual lengths (kinetic energy)
  const double* pedf;   // [3][NC] dual lengths of the flux (owner's for halo edges)
  const double* area;   // [6][NC]
  const double* hinv;   // [NC]
  const double* tri;    // [2][NC]
  const double* cor;    // [2][NC]
  const double* topo;   // [NC]
};

struct StageArgs {
  int n, S, L, ntile;
  int64_t NC;
  int64_t spole, npole;
  const double *ym, *yt, *yu;  // state the trend is evaluated on (ghost filled)
  const double *qm, *qt, *qu;  // base state of the update
  double *om, *ot, *ou;        // output state (or trend)
  double h;                    // update coefficient
  int trend_only;              // 1: out = trend
  Geo g;
  double grav, rho0, p_top;
  // optional copies of the intermediate quantities that the coarser levels
  // of a hierarchy restrict (null: not kept); FP64 pair kernel only
  double *k_hm, *k_ht;     // [L][3][NC] mass / temperature fluxes, region cells [-1, n]^2
  double *k_bern, *k_exn;  // [L][NC] Bernoulli, Exner on [0, n]^2
  double *k_src;           // [L][3][NC] velocity source (-Qperp) on owned edges
  double *k_dm, *k_dt;     // [L][NC] scalar trends on owned cells and the pole masters
  // split-explicit stage (scalar_star / u_star fused into the update): the
  // velocity update adds theta1 g grad eta to the trend (eta [NC], per member)
  const double* eta;
  double th1;
  // diffusion hooks (pair kernel, DIFF): scalar diffusive fluxes on the edges
  // of the region cells and velocity source on the owned edges, [L][3][NC]
  // (null: that hook is off)
  const double *phm, *pht, *phu;
};

__device__ __forceinline__ double interp(double a, double b) { return 0.5 * (a + b); }

// ensemble member blockIdx.y: its layers follow the previous member's in every
// state plane (scalars [member][layer][cell], velocity [member][layer][edge][cell])
__device__ __forceinline__ StageArgs member_args(const StageArgs& a0) {
  StageArgs a = a0;
  const int64_t s1 = (int64_t)blockIdx.y * a0.L * a0.NC, s3 = 3 * s1;
  a.ym += s1; a.yt += s1; a.yu += s3;
  a.qm += s1; a.qt += s1; a.qu += s3;
  a.om += s1; a.ot += s1; a.ou += s3;
  a.g.topo += (int64_t)blockIdx.y * a0.NC;  // each member has its own bathymetry
  if (a.eta) a.eta += (int64_t)blockIdx.y * a0.NC;
  if (a.phm) a.phm += s3, a.pht += s3;
  if (a.phu) a.phu += s3;
  return a;
}

// LDS working set of one block (15 planes of the 19 x 19 region)
struct Smem {
  double m[RR], t[RR], u[3][RR];
  double pvL[RR], pvU[RR];
  double qe[3][RR], hm[3][RR];
  double bern[RR], exn[RR];
};

// get_weights (ops.f90:1006-1021)
__device__ __forceinline__ void weights(const Geo& g, int64_t NC, int64_t c, int offs, double w[5]) {
  double acc0 = g.area[(int64_t)offs * NC + c];
  double acc1 = acc0 + g.area[(int64_t)((1 + offs) % 6) * NC + c];
  double acc2 = acc1 + g.area[(int64_t)((2 + offs) % 6) * NC + c];
  double acc3 = acc2 + g.area[(int64_t)((3 + offs) % 6) * NC + c];
  double acc4 = acc3 + g.area[(int64_t)((4 + offs) % 6) * NC + c];
  double hi = g.hinv[c];
  w[0] = 0.5 - acc0 * hi;
  w[1] = -(0.5 - acc1 * hi);
  w[2] = 0.5 - acc2 * hi;
  w[3] = -(0.5 - acc3 * hi);
  w[4] = 0.5 - acc4 * hi;
}

__global__ __launch_bounds__(THREADS) void trisk_stage_kernel(StageArgs a_in) {
  const StageArgs a = member_args(a_in);
  __shared__ Smem sm;
  const int tid = threadIdx.x;
  const int64_t NC = a.NC;
  const int S = a.S;
  const Geo& g = a.g;

  // XCD-aware tile order: block b runs on XCD b % 8, so give every XCD a
  // contiguous band of tiles (its halos then hit that XCD's L2)
  const int nb = gridDim.x;
  int b = blockIdx.x;
  int tile = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int tpl = a.ntile * a.ntile;
  const int lz = tile / tpl;
  const int rem = tile % tpl;
  const int i0 = (rem % a.ntile) * T, j0 = (rem / a.ntile) * T;
  // global cell of region cell (ri, rj) = lozenge (i0-1+ri, j0-1+rj)
  const int64_t base = (int64_t)lz * S * S + (int64_t)(j0 - 1 + H) * S + (i0 - 1 + H);
  auto gcell = [&](int ri, int rj) -> int64_t { return base + (int64_t)rj * S + ri; };
  auto r = [](int ri, int rj) { return rj * R + ri; };

  // --- column state: surface pressure and geopotential of the P cells ------
  // (cal_surf_press, set_surf_geopot; ops.f90:1023-1092)
  double pl[2], gp[2], pr[2], gpl[2];
  int pcell[2];
  for (int q = 0; q < 2; ++q) {
    int c = tid + q * THREADS;
    pcell[q] = c < NP * NP ? c : -1;
    pl[q] = gp[q] = pr[q] = gpl[q] = 0.0;
    if (pcell[q] < 0) continue;
    int ri = 1 + c % NP, rj = 1 + c / NP;
    int64_t gc = gcell(ri, rj);
    double s = 0.0;
    for (int k = 0; k < a.L; ++k) s = s + (a.ym[(int64_t)k * NC + gc] - a.yt[(int64_t)k * NC + gc]);
    pl[q] = a.grav * s + a.p_top;
    gp[q] = a.grav * g.topo[gc];
  }

  const bool cSW = (i0 == 0 && j0 == 0), cSE = (i0 + T == a.n && j0 == 0);
  const bool cNW = (i0 == 0 && j0 + T == a.n), cNE = (i0 + T == a.n && j0 + T == a.n);
  const bool corner = cSW || cSE || cNW || cNE;

  for (int k = 0; k < a.L; ++k) {
    const double* ym = a.ym + (int64_t)k * NC;
    const double* yt = a.yt + (int64_t)k * NC;
    const double* yu = a.yu + (int64_t)k * 3 * NC;
    // --- 1. stage the region; integrate pressure upward (ops.f90:1094-1131)
    for (int c = tid; c < RR; c += THREADS) {
      int64_t gc = gcell(c % R, c / R);
      sm.m[c] = ym[gc];
      sm.t[c] = yt[gc];
      sm.u[0][c] = yu[gc];
      sm.u[1][c] = yu[NC + gc];
      sm.u[2][c] = yu[2 * NC + gc];
    }
    for (int q = 0; q < 2; ++q) {
      if (pcell[q] < 0) continue;
      int64_t gc = gcell(1 + pcell[q] % NP, 1 + pcell[q] / NP);
      double rho_dz = ym[gc], rho_dz_theta = yt[gc];
      gpl[q] = gp[q];
      double p_upper = pl[q] - a.grav * (rho_dz - rho_dz_theta);
      pr[q] = interp(pl[q], p_upper);
      gp[q] = gpl[q] + a.grav * rho_dz / (a.rho0 * 1.0);
      pl[q] = p_upper;
    }
    __syncthreads();

    // --- 2. PV on triangles and mass fluxes on edges, cells -1 .. T ---------
    for (int c = tid; c < NT * NT; c += THREADS) {
      int ri = c % NT, rj = c / NT;
      int x = r(ri, rj), xE = r(ri + 1, rj), xN = r(ri, rj + 1), xNE = r(ri + 1, rj + 1);
      int64_t gc = gcell(ri, rj), gE = gcell(ri + 1, rj), gN = gcell(ri, rj + 1), gNE = gcell(ri + 1, rj + 1);
      double upRT = sm.u[0][x] * g.len[gc], upDG = sm.u[1][x] * g.len[NC + gc], upUP = sm.u[2][x] * g.len[2 * NC + gc];
      double circL = upRT + sm.u[2][xE] * g.len[2 * NC + gE] + upDG;
      double circU = -(upDG + upUP + sm.u[0][xN] * g.len[gN]);
      double massL = sm.m[x] * g.area[gc] + sm.m[xE] * g.area[2 * NC + gE] + sm.m[xNE] * g.area[4 * NC + gNE];
      double massU = sm.m[x] * g.area[NC + gc] + sm.m[xNE] * g.area[3 * NC + gNE] + sm.m[xN] * g.area[5 * NC + gN];
      sm.pvL[x] = (g.cor[gc] + circL) / massL;
      sm.pvU[x] = (g.cor[NC + gc] + circU) / massU;
      sm.hm[0][x] = sm.u[0][x] * g.pedf[gc] * interp(sm.m[x], sm.m[xE]) + 0.0;
      sm.hm[1][x] = sm.u[1][x] * g.pedf[NC + gc] * interp(sm.m[x], sm.m[xNE]) + 0.0;
      sm.hm[2][x] = sm.u[2][x] * g.pedf[2 * NC + gc] * interp(sm.m[x], sm.m[xN]) + 0.0;
    }
    __syncthreads();

    // --- 3. qe on edges; kinetic energy, Bernoulli, Exner on nodes 0 .. T ---
    for (int c = tid; c < NT * NT; c += THREADS) {
      int ri = c % NT, rj = c / NT, x = r(ri, rj);
      if (rj >= 1) sm.qe[0][x] = interp(sm.pvU[r(ri, rj - 1)], sm.pvL[x]);
      sm.qe[1][x] = interp(sm.pvU[x], sm.pvL[x]);
      if (ri >= 1) sm.qe[2][x] = interp(sm.pvU[x], sm.pvL[r(ri - 1, rj)]);
    }
    for (int q = 0; q < 2; ++q) {
      if (pcell[q] < 0) continue;
      int ri = 1 + pcell[q] % NP, rj = 1 + pcell[q] / NP;
      int x = r(ri, rj), xW = r(ri - 1, rj), xS = r(ri, rj - 1), xSW = r(ri - 1, rj - 1);
      int64_t gc = gcell(ri, rj), gW = gcell(ri - 1, rj), gS = gcell(ri, rj - 1), gSW = gcell(ri - 1, rj - 1);
      double ke = ((sm.u[2][x] * g.len[2 * NC + gc]) * (sm.u[2][x] * g.ped[2 * NC + gc]) +
                   (sm.u[1][x] * g.len[NC + gc]) * (sm.u[1][x] * g.ped[NC + gc]) +
                   (sm.u[0][x] * g.len[gc]) * (sm.u[0][x] * g.ped[gc]) +
                   (sm.u[2][xS] * g.len[2 * NC + gS]) * (sm.u[2][xS] * g.ped[2 * NC + gS]) +
                   (sm.u[1][xSW] * g.len[NC + gSW]) * (sm.u[1][xSW] * g.ped[NC + gSW]) +
                   (sm.u[0][xW] * g.len[gW]) * (sm.u[0][xW] * g.ped[gW])) *
                  g.hinv[gc] / 4.0;
      double phi_k = interp(gp[q], gpl[q]);
      sm.bern[x] = ke + phi_k + pr[q] / (a.rho0 * 1.0);
      sm.exn[x] = -phi_k;
    }
    __syncthreads();

    // --- 3b. pentagon corners (post_step1, ops.f90:526-737) -----------------
    if (corner) {
      if (tid < 4) {
        auto UPL = [&](int e, int ri, int rj) { return sm.u[e][r(ri, rj)] * g.len[(int64_t)e * NC + gcell(ri, rj)]; };
        auto M = [&](int ri, int rj) { return sm.m[r(ri, rj)]; };
        auto AR = [&](int p, int ri, int rj) { return g.area[(int64_t)(p - 1) * NC + gcell(ri, rj)]; };
        auto TRI = [&](int t, int ri, int rj) { return g.tri[(int64_t)t * NC + gcell(ri, rj)]; };
        auto COR = [&](int t, int ri, int rj) { return g.cor[(int64_t)t * NC + gcell(ri, rj)]; };
        if (tid == 0 && cSW) {  // IJMINUS at lozenge (0,0) = region (1,1)
          const int i = 1, j = 1;
          double cSW_L = UPL(2, i, j - 1) - UPL(0, i - 1, j) + UPL(0, i - 1, j - 1);
          double cW_L = ((UPL(0, i - 1, j) + UPL(2, i, j) + UPL(1, i - 1, j)) / TRI(0, i - 1, j)) * TRI(0, i - 1, j);
          double cS_U = (-(UPL(0, i, j) + UPL(1, i, j - 1) + UPL(2, i, j - 1)) / TRI(1, i, j - 1)) * TRI(1, i, j - 1);
          double m1 = M(i - 1, j) * AR(6, i - 1, j) + M(i, j) * (AR(4, i, j) + AR(5, i, j)) + M(i, j - 1) * AR(3, i, j - 1);
          double m2 = M(i - 1, j) * AR(1, i - 1, j) + M(i, j) * AR(3, i, j) + M(i, j + 1) * AR(5, i, j + 1);
          double m3 = M(i, j - 1) * AR(2, i, j - 1) + M(i + 1, j) * AR(4, i + 1, j) + M(i, j) * AR(6, i, j);
          double pvWL = (COR(0, i - 1, j) + cW_L) / m1;
          double pvSWL = (COR(0, i - 1, j - 1) + cSW_L) / m2;
          double pvSU = (COR(1, i, j - 1) + cS_U) / m3;
          sm.qe[0][r(i - 1, j)] = interp(pvWL, pvSWL);
          sm.qe[2][r(i, j - 1)] = interp(pvSU, pvSWL);
        }
        if (tid == 1 && cSE) {  // IPLUSJMINUS at lozenge (n,0) = region (T+1,1)
          const int i = T + 1, j = 1;
          double cSW_L = -UPL(0, i, j) + UPL(0, i - 1, j - 1) + UPL(1, i - 1, j - 1);
          double cL = ((UPL(0, i, j) + UPL(2, i + 1, j) + UPL(1, i, j)) / TRI(0, i, j)) * TRI(0, i, j);
          double cSW_U = (-(UPL(0, i - 1, j) + UPL(1, i - 1, j - 1) + UPL(2, i - 1, j - 1)) / TRI(1, i - 1, j - 1)) *
                         TRI(1, i - 1, j - 1);
          double m1 = M(i - 1, j - 1) * AR(1, i - 1, j - 1) + M(i, j - 1) * AR(3, i, j - 1) +
                      M(i, j) * (AR(5, i, j) + AR(6, i, j));
          double m2 = M(i, j) * AR(1, i, j) + M(i + 1, j) * AR(3, i + 1, j) + M(i + 1, j + 1) * AR(5, i + 1, j + 1);
          double m3 = M(i - 1, j - 1) * AR(2, i - 1, j - 1) + M(i, j) * AR(4, i, j) + M(i - 1, j) * AR(6, i - 1, j);
          double pvSWL = (COR(0, i - 1, j - 1) + cSW_L) / m1;
          double pvL = (COR(0, i, j) + cL) / m2;
          double pvSWU = (COR(1, i - 1, j - 1) + cSW_U) / m3;
          sm.qe[0][r(i, j)] = interp(pvSWL, pvL);
          sm.qe[1][r(i - 1, j - 1)] = interp(pvSWL, pvSWU);
        }
        if (tid == 2 && cNW) {  // IMINUSJPLUS at lozenge (0,n) = region (1,T+1)
          const int i = 1, j = T + 1;
          double cSW_U = UPL(2, i, j) - UPL(1, i - 1, j - 1) - UPL(2, i - 1, j - 1);
          double cU = (-(UPL(1, i, j) + UPL(2, i, j) + UPL(0, i, j + 1)) / TRI(1, i, j)) * TRI(1, i, j);
          double cSW_L = ((UPL(0, i - 1, j - 1) + UPL(2, i, j - 1) + UPL(1, i - 1, j - 1)) / TRI(0, i - 1, j - 1)) *
                         TRI(0, i - 1, j - 1);
          double m1 = M(i - 1, j - 1) * AR(2, i - 1, j - 1) + M(i, j) * (AR(3, i, j) + AR(4, i, j)) +
                      M(i - 1, j) * AR(6, i - 1, j);
          double m2 = M(i, j) * AR(2, i, j) + M(i + 1, j + 1) * AR(4, i + 1, j + 1) + M(i, j + 1) * AR(6, i, j + 1);
          double m3 = M(i - 1, j - 1) * AR(1, i - 1, j - 1) + M(i, j - 1) * AR(3, i, j - 1) + M(i, j) * AR(5, i, j);
          double pvSWU = (COR(1, i - 1, j - 1) + cSW_U) / m1;
          double pvU = (COR(1, i, j) + cU) / m2;
          double pvSWL = (COR(0, i - 1, j - 1) + cSW_L) / m3;
          sm.qe[2][r(i, j)] = interp(pvSWU, pvU);
          sm.qe[1][r(i - 1, j - 1)] = interp(pvSWL, pvSWU);
        }
        if (tid == 3 && cNE) {  // IJPLUS at lozenge (n,n) = region (T+1,T+1)
          const int i = T + 1, j = T + 1;
          double cL = UPL(0, i, j) - UPL(0, i, j + 1) - UPL(2, i, j);
          double cW_L = ((UPL(0, i - 1, j) + UPL(2, i, j) + UPL(1, i - 1, j)) / TRI(0, i - 1, j)) * TRI(0, i - 1, j);
          double cS_U = (-(UPL(1, i, j - 1) + UPL(2, i, j - 1) + UPL(0, i, j)) / TRI(1, i, j - 1)) * TRI(1, i, j - 1);
          double m1 = M(i + 1, j) * AR(3, i + 1, j) + M(i, j) * (AR(1, i, j) + AR(2, i, j)) + M(i, j + 1) * AR(6, i, j + 1);
          double m2 = M(i - 1, j) * AR(1, i - 1, j) + M(i, j) * AR(3, i, j) + M(i, j + 1) * AR(5, i, j + 1);
          double m3 = M(i, j - 1) * AR(2, i, j - 1) + M(i + 1, j) * AR(4, i + 1, j) + M(i, j) * AR(6, i, j);
          double pvL = (COR(0, i, j) + cL) / m1;
          double pvWL = (COR(0, i - 1, j) + cW_L) / m2;
          double pvSU = (COR(1, i, j - 1) + cS_U) / m3;
          sm.qe[0][r(i, j)] = interp(pvL, pvSU);
          sm.qe[2][r(i, j)] = interp(pvL, pvWL);
        }
      }
      __syncthreads();
    }

    // --- 4. trends of owned nodes and edges; fused RK update -----------------
    {
      const int ri = 1 + tid % T, rj = 1 + tid / T;
      const int x = r(ri, rj), xW = r(ri - 1, rj), xS = r(ri, rj - 1), xSW = r(ri - 1, rj - 1);
      const int xE = r(ri + 1, rj), xN = r(ri, rj + 1), xNE = r(ri + 1, rj + 1), xNW = r(ri - 1, rj + 1);
      const int xSE = r(ri + 1, rj - 1);
      const int64_t gc = gcell(ri, rj);
      const double hi = g.hinv[gc];
      // scalar trends (scalar_trend / div, ops.f90:739-755, 1325-1343)
      double divm = (sm.hm[0][x] - sm.hm[0][xW] + sm.hm[1][xSW] - sm.hm[1][x] + sm.hm[2][x] - sm.hm[2][xS]) * hi;
      auto ft = [&](int e, int xa, int xb, int64_t ga) {
        return sm.u[e][xa] * g.pedf[(int64_t)e * NC + ga] * interp(sm.t[xa], sm.t[xb]) + 0.0;
      };
      const int64_t gW = gcell(ri - 1, rj), gS = gcell(ri, rj - 1), gSW = gcell(ri - 1, rj - 1);
      double divt = (ft(0, x, xE, gc) - ft(0, xW, x, gW) + ft(1, xSW, x, gSW) - ft(1, x, xNE, gc) + ft(2, x, xN, gc) -
                     ft(2, xS, x, gS)) *
                    hi;
      double dmv = -divm, dtv = -divt;
      // velocity trend: -Qperp (ops.f90:834-901), then gradients (ops.f90:787-832)
      const int64_t gE = gcell(ri + 1, rj), gNE = gcell(ri + 1, rj + 1), gN = gcell(ri, rj + 1);
      double w1[5], w2[5], qp[3];
#define HM(e, xx) sm.hm[e][xx]
#define QE(e, xx) sm.qe[e][xx]
      weights(g, NC, gc, 0, w1);
      weights(g, NC, gE, 3, w2);
      double qr = QE(0, x);
      qp[0] = HM(1, x) * interp(QE(1, x), qr) * w1[0] + HM(2, x) * interp(QE(2, x), qr) * w1[1] +
              HM(0, xW) * interp(QE(0, xW), qr) * w1[2] + HM(1, xSW) * interp(QE(1, xSW), qr) * w1[3] +
              HM(2, xS) * interp(QE(2, xS), qr) * w1[4] + HM(1, xS) * interp(QE(1, xS), qr) * w2[0] +
              HM(2, xSE) * interp(QE(2, xSE), qr) * w2[1] + HM(0, xE) * interp(QE(0, xE), qr) * w2[2] +
              HM(1, xE) * interp(QE(1, xE), qr) * w2[3] + HM(2, xE) * interp(QE(2, xE), qr) * w2[4];
      weights(g, NC, gc, 1, w1);
      weights(g, NC, gNE, 4, w2);
      double qd = QE(1, x);
      qp[1] = HM(2, x) * interp(QE(2, x), qd) * w1[0] + HM(0, xW) * interp(QE(0, xW), qd) * w1[1] +
              HM(1, xSW) * interp(QE(1, xSW), qd) * w1[2] + HM(2, xS) * interp(QE(2, xS), qd) * w1[3] +
              HM(0, x) * interp(QE(0, x), qd) * w1[4] + HM(2, xE) * interp(QE(2, xE), qd) * w2[0] +
              HM(0, xNE) * interp(QE(0, xNE), qd) * w2[1] + HM(1, xNE) * interp(QE(1, xNE), qd) * w2[2] +
              HM(2, xNE) * interp(QE(2, xNE), qd) * w2[3] + HM(0, xN) * interp(QE(0, xN), qd) * w2[4];
      weights(g, NC, gc, 2, w1);
      weights(g, NC, gN, 5, w2);
      double qu = QE(2, x);
      qp[2] = HM(0, xW) * interp(QE(0, xW), qu) * w1[0] + HM(1, xSW) * interp(QE(1, xSW), qu) * w1[1] +
              HM(2, xS) * interp(QE(2, xS), qu) * w1[2] + HM(0, x) * interp(QE(0, x), qu) * w1[3] +
              HM(1, x) * interp(QE(1, x), qu) * w1[4] + HM(0, xN) * interp(QE(0, xN), qu) * w2[0] +
              HM(1, xN) * interp(QE(1, xN), qu) * w2[1] + HM(2, xN) * interp(QE(2, xN), qu) * w2[2] +
              HM(0, xNW) * interp(QE(0, xNW), qu) * w2[3] + HM(1, xW) * interp(QE(1, xW), qu) * w2[4];
#undef HM
#undef QE
      double th0 = sm.t[x] / sm.m[x], thE = sm.t[xE] / sm.m[xE], thNE = sm.t[xNE] / sm.m[xNE],
             thN = sm.t[xN] / sm.m[xN];
      double the[3] = {interp(th0, thE), interp(th0, thNE), interp(th0, thN)};
      double l0 = g.len[gc], l1 = g.len[NC + gc], l2 = g.len[2 * NC + gc];
      double gB[3] = {(sm.bern[xE] - sm.bern[x]) / l0, (sm.bern[x] - sm.bern[xNE]) / l1, (sm.bern[xN] - sm.bern[x]) / l2};
      double gX[3] = {(sm.exn[xE] - sm.exn[x]) / l0, (sm.exn[x] - sm.exn[xNE]) / l1, (sm.exn[xN] - sm.exn[x]) / l2};
      double ll[3] = {l0, l1, l2};
      const int64_t ks = (int64_t)k * NC, ke3 = (int64_t)k * 3 * NC;
      if (a.trend_only) {
        a.om[ks + gc] = dmv;
        a.ot[ks + gc] = dtv;
      } else {
        a.om[ks + gc] = a.qm[ks + gc] + a.h * dmv;
        a.ot[ks + gc] = a.qt[ks + gc] + a.h * dtv;
      }
      for (int e = 0; e < 3; ++e) {
        double src = -qp[e] + 0.0 * ll[e];
        double du = src / ll[e] - gB[e] - the[e] * gX[e];
        int64_t ge = ke3 + (int64_t)e * NC + gc;
        a.ou[ge] = a.trend_only ? du : a.qu[ge] + a.h * du;
      }
      // pole master cell (south: lozenge 0 cell (n,0); north: lozenge 5 (0,n))
      if (tid == 0 && ((lz == 0 && cSE) || (lz == 5 && cNW))) {
        const int pi = (lz == 0) ? T + 1 : 1, pj = (lz == 0) ? 1 : T + 1;
        const int y = r(pi, pj), yW = r(pi - 1, pj), yS = r(pi, pj - 1), ySW = r(pi - 1, pj - 1);
        const int yE = r(pi + 1, pj), yN = r(pi, pj + 1), yNE = r(pi + 1, pj + 1);
        const int64_t pc = gcell(pi, pj);
        const double phi = g.hinv[pc];
        double pdm = -((sm.hm[0][y] - sm.hm[0][yW] + sm.hm[1][ySW] - sm.hm[1][y] + sm.hm[2][y] - sm.hm[2][yS]) * phi);
        const int64_t pW = gcell(pi - 1, pj), pS = gcell(pi, pj - 1), pSW = gcell(pi - 1, pj - 1);
        double pdt = -((ft(0, y, yE, pc) - ft(0, yW, y, pW) + ft(1, ySW, y, pSW) - ft(1, y, yNE, pc) +
                        ft(2, y, yN, pc) - ft(2, yS, y, pS)) *
                       phi);
        if (a.trend_only) {
          a.om[ks + pc] = pdm;
          a.ot[ks + pc] = pdt;
        } else {
          a.om[ks + pc] = a.qm[ks + pc] + a.h * pdm;
          a.ot[ks + pc] = a.qt[ks + pc] + a.h * pdt;
        }
      }
    }
    __syncthreads();
  }
}

// post_step1 (ops.f90:526-737): at the four lozenge corners the PV of the
// triangles next to the pentagon is recomputed with the pentagon's
// circulation and mass, and two qe values per corner are overwritten.
// `which` = 0..3 for the SW, SE, NW, NE corner; region cell (1,1), (T+1,1),
// (1,T+1), (T+1,T+1).  u*, m, qe* are LDS planes of the 19 x 19 region.
// The planes are FP64 or FP32 (F); the corrections are evaluated in FP64.
template <class F>
__device__ void pentagon_fixup(int which, const F* __restrict__ u0p, const F* __restrict__ u1p,
                               const F* __restrict__ u2p, const F* __restrict__ mp, F* qe0, F* qe1, F* qe2,
                               const Geo& g, int64_t NC, int64_t base, int S) {
  auto r = [](int ri, int rj) { return rj * R + ri; };
  auto gcell = [&](int ri, int rj) -> int64_t { return base + (int64_t)rj * S + ri; };
  const F* up[3] = {u0p, u1p, u2p};
  auto UPL = [&](int e, int i, int j) { return (double)up[e][r(i, j)] * g.len[(int64_t)e * NC + gcell(i, j)]; };
  auto M = [&](int i, int j) { return (double)mp[r(i, j)]; };
  auto AR = [&](int p, int i, int j) { return g.area[(int64_t)(p - 1) * NC + gcell(i, j)]; };
  auto TRI = [&](int tt, int i, int j) { return g.tri[(int64_t)tt * NC + gcell(i, j)]; };
  auto COR = [&](int tt, int i, int j) { return g.cor[(int64_t)tt * NC + gcell(i, j)]; };
  if (which == 0) {  // IJMINUS
    const int i = 1, j = 1;
    double cSW_L = UPL(2, i, j - 1) - UPL(0, i - 1, j) + UPL(0, i - 1, j - 1);
    double cW_L = ((UPL(0, i - 1, j) + UPL(2, i, j) + UPL(1, i - 1, j)) / TRI(0, i - 1, j)) * TRI(0, i - 1, j);
    double cS_U = (-(UPL(0, i, j) + UPL(1, i, j - 1) + UPL(2, i, j - 1)) / TRI(1, i, j - 1)) * TRI(1, i, j - 1);
    double m1 = M(i - 1, j) * AR(6, i - 1, j) + M(i, j) * (AR(4, i, j) + AR(5, i, j)) + M(i, j - 1) * AR(3, i, j - 1);
    double m2 = M(i - 1, j) * AR(1, i - 1, j) + M(i, j) * AR(3, i, j) + M(i, j + 1) * AR(5, i, j + 1);
    double m3 = M(i, j - 1) * AR(2, i, j - 1) + M(i + 1, j) * AR(4, i + 1, j) + M(i, j) * AR(6, i, j);
    double pvWL = (COR(0, i - 1, j) + cW_L) / m1;
    double pvSWL = (COR(0, i - 1, j - 1) + cSW_L) / m2;
    double pvSU = (COR(1, i, j - 1) + cS_U) / m3;
    qe0[r(i - 1, j)] = (F)interp(pvWL, pvSWL);
    qe2[r(i, j - 1)] = (F)interp(pvSU, pvSWL);
  } else if (which == 1) {  // IPLUSJMINUS
    const int i = T + 1, j = 1;
    double cSW_L = -UPL(0, i, j) + UPL(0, i - 1, j - 1) + UPL(1, i - 1, j - 1);
    double cL = ((UPL(0, i, j) + UPL(2, i + 1, j) + UPL(1, i, j)) / TRI(0, i, j)) * TRI(0, i, j);
    double cSW_U = (-(UPL(0, i - 1, j) + UPL(1, i - 1, j - 1) + UPL(2, i - 1, j - 1)) / TRI(1, i - 1, j - 1)) *
                   TRI(1, i - 1, j - 1);
    double m1 =
        M(i - 1, j - 1) * AR(1, i - 1, j - 1) + M(i, j - 1) * AR(3, i, j - 1) + M(i, j) * (AR(5, i, j) + AR(6, i, j));
    double m2 = M(i, j) * AR(1, i, j) + M(i + 1, j) * AR(3, i + 1, j) + M(i + 1, j + 1) * AR(5, i + 1, j + 1);
    double m3 = M(i - 1, j - 1) * AR(2, i - 1, j - 1) + M(i, j) * AR(4, i, j) + M(i - 1, j) * AR(6, i - 1, j);
    double pvSWL = (COR(0, i - 1, j - 1) + cSW_L) / m1;
    double pvL = (COR(0, i, j) + cL) / m2;
    double pvSWU = (COR(1, i - 1, j - 1) + cSW_U) / m3;
    qe0[r(i, j)] = (F)interp(pvSWL, pvL);
    qe1[r(i - 1, j - 1)] = (F)interp(pvSWL, pvSWU);
  } else if (which == 2) {  // IMINUSJPLUS
    const int i = 1, j = T + 1;
    double cSW_U = UPL(2, i, j) - UPL(1, i - 1, j - 1) - UPL(2, i - 1, j - 1);
    double cU = (-(UPL(1, i, j) + UPL(2, i, j) + UPL(0, i, j + 1)) / TRI(1, i, j)) * TRI(1, i, j);
    double cSW_L = ((UPL(0, i - 1, j - 1) + UPL(2, i, j - 1) + UPL(1, i - 1, j - 1)) / TRI(0, i - 1, j - 1)) *
                   TRI(0, i - 1, j - 1);
    double m1 =
        M(i - 1, j - 1) * AR(2, i - 1, j - 1) + M(i, j) * (AR(3, i, j) + AR(4, i, j)) + M(i - 1, j) * AR(6, i - 1, j);
    double m2 = M(i, j) * AR(2, i, j) + M(i + 1, j + 1) * AR(4, i + 1, j + 1) + M(i, j + 1) * AR(6, i, j + 1);
    double m3 = M(i - 1, j - 1) * AR(1, i - 1, j - 1) + M(i, j - 1) * AR(3, i, j - 1) + M(i, j) * AR(5, i, j);
    double pvSWU = (COR(1, i - 1, j - 1) + cSW_U) / m1;
    double pvU = (COR(1, i, j) + cU) / m2;
    double pvSWL = (COR(0, i - 1, j - 1) + cSW_L) / m3;
    qe2[r(i, j)] = (F)interp(pvSWU, pvU);
    qe1[r(i - 1, j - 1)] = (F)interp(pvSWL, pvSWU);
  } else {  // IJPLUS
    const int i = T + 1, j = T + 1;
    double cL = UPL(0, i, j) - UPL(0, i, j + 1) - UPL(2, i, j);
    double cW_L = ((UPL(0, i - 1, j) + UPL(2, i, j) + UPL(1, i - 1, j)) / TRI(0, i - 1, j)) * TRI(0, i - 1, j);
    double cS_U = (-(UPL(1, i, j - 1) + UPL(2, i, j - 1) + UPL(0, i, j)) / TRI(1, i, j - 1)) * TRI(1, i, j - 1);
    double m1 = M(i + 1, j) * AR(3, i + 1, j) + M(i, j) * (AR(1, i, j) + AR(2, i, j)) + M(i, j + 1) * AR(6, i, j + 1);
    double m2 = M(i - 1, j) * AR(1, i - 1, j) + M(i, j) * AR(3, i, j) + M(i, j + 1) * AR(5, i, j + 1);
    double m3 = M(i, j - 1) * AR(2, i, j - 1) + M(i + 1, j) * AR(4, i + 1, j) + M(i, j) * AR(6, i, j);
    double pvL = (COR(0, i, j) + cL) / m1;
    double pvWL = (COR(0, i - 1, j) + cW_L) / m2;
    double pvSU = (COR(1, i, j - 1) + cS_U) / m3;
    qe0[r(i, j)] = (F)interp(pvL, pvSU);
    qe2[r(i, j)] = (F)interp(pvL, pvWL);
  }
}

// Loads and stores at a uniform base pointer plus a 32-bit byte offset, so the
// per-thread part of an address is one VGPR (global_* with saddr).
__device__ __forceinline__ double ldo(const double* p, uint32_t off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) + off);
}
__device__ __forceinline__ void sto(double* p, uint32_t off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(p) + off) = v;
}

// sum over layers of (mass - temp) of one column, in layer order (the first
// step of cal_surf_press, ops.f90:1023-1049); the loads of a chunk of
// layers are issued together so the column costs a few memory round trips
__device__ __forceinline__ double column_sum(const double* ym, const double* yt, int L, int64_t NC, uint32_t xb) {
  constexpr int CH = 10;
  double s = 0.0;
  for (int k0 = 0; k0 < L; k0 += CH) {
    double dm[CH], dt[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int k = k0 + j < L ? k0 + j : L - 1;  // clamped: in bounds, not summed
      dm[j] = ldo(ym + k * NC, xb);
      dt[j] = ldo(yt + k * NC, xb);
    }
    __builtin_amdgcn_sched_barrier(0);  // all loads of the chunk before the first add
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (k0 + j < L) s = s + (dm[j] - dt[j]);
  }
  return s;
}

// ---------------------------------------------------------------------------
// trisk_stage_kernel2: same arithmetic as trisk_stage_kernel, organised for
// throughput.  One thread per cell of the 19 x 19 staged region (384 threads,
// 361 active), fixed for the whole launch, so every layer-invariant
// geometric coefficient a cell needs -- its own lengths, areas and Coriolis
// terms and the neighbour-derived ones (the Qperp weights of E, NE and N,
// the area parts of the triangle vertices, the lengths of the W/SW/S edges
// of its hexagon) -- is loaded once into registers before the layer loop.
// Per layer only the state is read from HBM; scalar fluxes of both tracers
// go through LDS so the owned cells read them instead of recomputing.
// ---------------------------------------------------------------------------
constexpr int THREADS2 = 384;

struct Smem2 {
  double m[RR], t[RR], u[3][RR];
  double pvL[RR], pvU[RR];
  double qe[3][RR], hm[3][RR], ht[3][RR];
  double bern[RR], exn[RR];
};

__device__ __forceinline__ void weights_from(double a1, double a2, double a3, double a4, double a5, double a6,
                                             double hi, int offs, double w[5]) {
  const double a[6] = {a1, a2, a3, a4, a5, a6};
  double acc0 = a[offs];
  double acc1 = acc0 + a[(1 + offs) % 6];
  double acc2 = acc1 + a[(2 + offs) % 6];
  double acc3 = acc2 + a[(3 + offs) % 6];
  double acc4 = acc3 + a[(4 + offs) % 6];
  w[0] = 0.5 - acc0 * hi;
  w[1] = -(0.5 - acc1 * hi);
  w[2] = 0.5 - acc2 * hi;
  w[3] = -(0.5 - acc3 * hi);
  w[4] = 0.5 - acc4 * hi;
}

__global__ __launch_bounds__(THREADS2, 2) void trisk_stage_kernel2(StageArgs a_in) {
  const StageArgs a = member_args(a_in);
  __shared__ Smem2 sm;
  const int tid = threadIdx.x;
  const int64_t NC = a.NC;
  const int S = a.S;
  const Geo& g = a.g;

  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int tile = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int tpl = a.ntile * a.ntile;
  const int lz = tile / tpl;
  const int rem = tile % tpl;
  const int i0 = (rem % a.ntile) * T, j0 = (rem / a.ntile) * T;
  const int64_t base = (int64_t)lz * S * S + (int64_t)(j0 - 1 + H) * S + (i0 - 1 + H);
  auto gcell = [&](int ri, int rj) -> int64_t { return base + (int64_t)rj * S + ri; };
  auto r = [](int ri, int rj) { return rj * R + ri; };

  const bool cSW = (i0 == 0 && j0 == 0), cSE = (i0 + T == a.n && j0 == 0);
  const bool cNW = (i0 == 0 && j0 + T == a.n), cNE = (i0 + T == a.n && j0 + T == a.n);
  const bool corner = cSW || cSE || cNW || cNE;

  // this thread's cell and roles
  const bool active = tid < RR;
  const int ri = tid % R, rj = tid / R;
  const int x = tid;
  const int64_t gc = active ? gcell(ri, rj) : base;
  const bool tri_cell = active && ri <= T + 1 && rj <= T + 1;                   // PV, fluxes, qe
  const bool p_cell = active && ri >= 1 && rj >= 1 && ri <= T + 1 && rj <= T + 1;  // pressure, KE, Bernoulli
  const bool own = active && ri >= 1 && rj >= 1 && ri <= T && rj <= T;
  const bool pole = (lz == 0 && cSE && ri == T + 1 && rj == 1) || (lz == 5 && cNW && ri == 1 && rj == T + 1);

  // --- layer-invariant geometry, once per launch ------------------------------
  double len0 = 0, len1 = 0, len2 = 0, ped0 = 0, ped1 = 0, ped2 = 0, pf0 = 0, pf1 = 0, pf2 = 0;
  double A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0, A6 = 0, hi = 0, corL = 0, corU = 0;
  double lenUPE = 0, lenRTN = 0, A3E = 0, A5NE = 0, A4NE = 0, A6N = 0;
  double lenRTW = 0, pedRTW = 0, lenDGSW = 0, pedDGSW = 0, lenUPS = 0, pedUPS = 0;
  double w2E[5] = {0, 0, 0, 0, 0}, w2NE[5] = {0, 0, 0, 0, 0}, w2N[5] = {0, 0, 0, 0, 0};
  if (tri_cell) {
    len0 = g.len[gc]; len1 = g.len[NC + gc]; len2 = g.len[2 * NC + gc];
    pf0 = g.pedf[gc]; pf1 = g.pedf[NC + gc]; pf2 = g.pedf[2 * NC + gc];
    A1 = g.area[gc]; A2 = g.area[NC + gc];
    corL = g.cor[gc]; corU = g.cor[NC + gc];
    const int64_t gE = gcell(ri + 1, rj), gN = gcell(ri, rj + 1), gNE = gcell(ri + 1, rj + 1);
    lenUPE = g.len[2 * NC + gE]; lenRTN = g.len[gN];
    A3E = g.area[2 * NC + gE]; A5NE = g.area[4 * NC + gNE]; A4NE = g.area[3 * NC + gNE]; A6N = g.area[5 * NC + gN];
  }
  if (p_cell) {
    ped0 = g.ped[gc]; ped1 = g.ped[NC + gc]; ped2 = g.ped[2 * NC + gc];
    A3 = g.area[2 * NC + gc]; A4 = g.area[3 * NC + gc]; A5 = g.area[4 * NC + gc]; A6 = g.area[5 * NC + gc];
    hi = g.hinv[gc];
    const int64_t gW = gcell(ri - 1, rj), gS = gcell(ri, rj - 1), gSW = gcell(ri - 1, rj - 1);
    lenRTW = g.len[gW]; pedRTW = g.ped[gW];
    lenDGSW = g.len[NC + gSW]; pedDGSW = g.ped[NC + gSW];
    lenUPS = g.len[2 * NC + gS]; pedUPS = g.ped[2 * NC + gS];
  }
  if (own) {
    weights(g, NC, gcell(ri + 1, rj), 3, w2E);
    weights(g, NC, gcell(ri + 1, rj + 1), 4, w2NE);
    weights(g, NC, gcell(ri, rj + 1), 5, w2N);
  }

  // column state (cal_surf_press, set_surf_geopot)
  double pl = 0, gp = 0, pr = 0, gpl = 0;
  if (p_cell) {
    double s = column_sum(a.ym, a.yt, a.L, NC, (uint32_t)gc * 8u);
    pl = a.grav * s + a.p_top;
    gp = a.grav * g.topo[gc];
  }

  // state of layer 0; later layers are prefetched one layer ahead
  double m = 0, t = 0, u0 = 0, u1 = 0, u2 = 0;
  if (active) {
    m = a.ym[gc];
    t = a.yt[gc];
    u0 = a.yu[gc];
    u1 = a.yu[NC + gc];
    u2 = a.yu[2 * NC + gc];
  }

  for (int k = 0; k < a.L; ++k) {
    const int64_t ks = (int64_t)k * NC, ke3 = (int64_t)k * 3 * NC;
    // --- 1. stage own cell; integrate pressure upward ------------------------
    if (active) {
      sm.m[x] = m;
      sm.t[x] = t;
      sm.u[0][x] = u0;
      sm.u[1][x] = u1;
      sm.u[2][x] = u2;
    }
    double nm = 0, nt = 0, nu0 = 0, nu1 = 0, nu2 = 0;
    if (active && k + 1 < a.L) {
      const int64_t ks1 = ks + NC, ke31 = ke3 + 3 * NC;
      nm = a.ym[ks1 + gc];
      nt = a.yt[ks1 + gc];
      nu0 = a.yu[ke31 + gc];
      nu1 = a.yu[ke31 + NC + gc];
      nu2 = a.yu[ke31 + 2 * NC + gc];
    }
    // base state of the update, consumed in phase 4
    double bm = 0, bt = 0, bu0 = 0, bu1 = 0, bu2 = 0;
    if (!a.trend_only) {
      if (own || pole) {
        bm = a.qm[ks + gc];
        bt = a.qt[ks + gc];
      }
      if (own) {
        bu0 = a.qu[ke3 + gc];
        bu1 = a.qu[ke3 + NC + gc];
        bu2 = a.qu[ke3 + 2 * NC + gc];
      }
    }
    if (p_cell) {
      gpl = gp;
      double p_upper = pl - a.grav * (m - t);
      pr = interp(pl, p_upper);
      gp = gpl + a.grav * m / (a.rho0 * 1.0);
      pl = p_upper;
    }
    __syncthreads();

    // --- 2. PV on the cell's two triangles; mass and tracer fluxes ------------
    if (tri_cell) {
      const int xE = x + 1, xN = x + R, xNE = x + R + 1;
      double upRT = u0 * len0, upDG = u1 * len1, upUP = u2 * len2;
      double circL = upRT + sm.u[2][xE] * lenUPE + upDG;
      double circU = -(upDG + upUP + sm.u[0][xN] * lenRTN);
      double mE = sm.m[xE], mNE = sm.m[xNE], mN = sm.m[xN];
      double massL = m * A1 + mE * A3E + mNE * A5NE;
      double massU = m * A2 + mNE * A4NE + mN * A6N;
      sm.pvL[x] = (corL + circL) / massL;
      sm.pvU[x] = (corU + circU) / massU;
      sm.hm[0][x] = u0 * pf0 * interp(m, mE) + 0.0;
      sm.hm[1][x] = u1 * pf1 * interp(m, mNE) + 0.0;
      sm.hm[2][x] = u2 * pf2 * interp(m, mN) + 0.0;
      sm.ht[0][x] = u0 * pf0 * interp(t, sm.t[xE]) + 0.0;
      sm.ht[1][x] = u1 * pf1 * interp(t, sm.t[xNE]) + 0.0;
      sm.ht[2][x] = u2 * pf2 * interp(t, sm.t[xN]) + 0.0;
    }
    __syncthreads();

    // --- 3. qe; kinetic energy, Bernoulli, Exner ------------------------------
    if (tri_cell) {
      if (rj >= 1) sm.qe[0][x] = interp(sm.pvU[x - R], sm.pvL[x]);
      sm.qe[1][x] = interp(sm.pvU[x], sm.pvL[x]);
      if (ri >= 1) sm.qe[2][x] = interp(sm.pvU[x], sm.pvL[x - 1]);
    }
    if (p_cell) {
      double uS = sm.u[2][x - R], uSW = sm.u[1][x - R - 1], uW = sm.u[0][x - 1];
      double ke = ((u2 * len2) * (u2 * ped2) + (u1 * len1) * (u1 * ped1) + (u0 * len0) * (u0 * ped0) +
                   (uS * lenUPS) * (uS * pedUPS) + (uSW * lenDGSW) * (uSW * pedDGSW) + (uW * lenRTW) * (uW * pedRTW)) *
                  hi / 4.0;
      double phi_k = interp(gp, gpl);
      sm.bern[x] = ke + phi_k + pr / (a.rho0 * 1.0);
      sm.exn[x] = -phi_k;
    }
    __syncthreads();

    // --- 3b. pentagon corners (post_step1) -------------------------------------
    if (corner) {
      const bool mine[4] = {cSW, cSE, cNW, cNE};
      if (tid < 4 && mine[tid])
        pentagon_fixup(tid, sm.u[0], sm.u[1], sm.u[2], sm.m, sm.qe[0], sm.qe[1], sm.qe[2], g, NC, base, S);
      __syncthreads();
    }

    // --- 4. trends of owned nodes/edges and the pole master; RK update --------
    if (own || pole) {
      const int xW = x - 1, xS = x - R, xSW = x - R - 1;
      double divm = (sm.hm[0][x] - sm.hm[0][xW] + sm.hm[1][xSW] - sm.hm[1][x] + sm.hm[2][x] - sm.hm[2][xS]) * hi;
      double divt = (sm.ht[0][x] - sm.ht[0][xW] + sm.ht[1][xSW] - sm.ht[1][x] + sm.ht[2][x] - sm.ht[2][xS]) * hi;
      double dmv = -divm, dtv = -divt;
      a.om[ks + gc] = a.trend_only ? dmv : bm + a.h * dmv;
      a.ot[ks + gc] = a.trend_only ? dtv : bt + a.h * dtv;
    }
    if (own) {
      const int xW = x - 1, xS = x - R, xSW = x - R - 1, xE = x + 1, xN = x + R, xNE = x + R + 1;
      const int xNW = x + R - 1, xSE = x - R + 1;
      double w1[5], qp[3];
#define HM(e, xx) sm.hm[e][xx]
#define QE(e, xx) sm.qe[e][xx]
      weights_from(A1, A2, A3, A4, A5, A6, hi, 0, w1);
      double qr = QE(0, x);
      qp[0] = HM(1, x) * interp(QE(1, x), qr) * w1[0] + HM(2, x) * interp(QE(2, x), qr) * w1[1] +
              HM(0, xW) * interp(QE(0, xW), qr) * w1[2] + HM(1, xSW) * interp(QE(1, xSW), qr) * w1[3] +
              HM(2, xS) * interp(QE(2, xS), qr) * w1[4] + HM(1, xS) * interp(QE(1, xS), qr) * w2E[0] +
              HM(2, xSE) * interp(QE(2, xSE), qr) * w2E[1] + HM(0, xE) * interp(QE(0, xE), qr) * w2E[2] +
              HM(1, xE) * interp(QE(1, xE), qr) * w2E[3] + HM(2, xE) * interp(QE(2, xE), qr) * w2E[4];
      weights_from(A1, A2, A3, A4, A5, A6, hi, 1, w1);
      double qd = QE(1, x);
      qp[1] = HM(2, x) * interp(QE(2, x), qd) * w1[0] + HM(0, xW) * interp(QE(0, xW), qd) * w1[1] +
              HM(1, xSW) * interp(QE(1, xSW), qd) * w1[2] + HM(2, xS) * interp(QE(2, xS), qd) * w1[3] +
              HM(0, x) * interp(QE(0, x), qd) * w1[4] + HM(2, xE) * interp(QE(2, xE), qd) * w2NE[0] +
              HM(0, xNE) * interp(QE(0, xNE), qd) * w2NE[1] + HM(1, xNE) * interp(QE(1, xNE), qd) * w2NE[2] +
              HM(2, xNE) * interp(QE(2, xNE), qd) * w2NE[3] + HM(0, xN) * interp(QE(0, xN), qd) * w2NE[4];
      weights_from(A1, A2, A3, A4, A5, A6, hi, 2, w1);
      double qu = QE(2, x);
      qp[2] = HM(0, xW) * interp(QE(0, xW), qu) * w1[0] + HM(1, xSW) * interp(QE(1, xSW), qu) * w1[1] +
              HM(2, xS) * interp(QE(2, xS), qu) * w1[2] + HM(0, x) * interp(QE(0, x), qu) * w1[3] +
              HM(1, x) * interp(QE(1, x), qu) * w1[4] + HM(0, xN) * interp(QE(0, xN), qu) * w2N[0] +
              HM(1, xN) * interp(QE(1, xN), qu) * w2N[1] + HM(2, xN) * interp(QE(2, xN), qu) * w2N[2] +
              HM(0, xNW) * interp(QE(0, xNW), qu) * w2N[3] + HM(1, xW) * interp(QE(1, xW), qu) * w2N[4];
#undef HM
#undef QE
      double th0 = t / m, thE = sm.t[xE] / sm.m[xE], thNE = sm.t[xNE] / sm.m[xNE], thN = sm.t[xN] / sm.m[xN];
      double the[3] = {interp(th0, thE), interp(th0, thNE), interp(th0, thN)};
      double b0 = sm.bern[x], x0 = sm.exn[x];
      double gB[3] = {(sm.bern[xE] - b0) / len0, (b0 - sm.bern[xNE]) / len1, (sm.bern[xN] - b0) / len2};
      double gX[3] = {(sm.exn[xE] - x0) / len0, (x0 - sm.exn[xNE]) / len1, (sm.exn[xN] - x0) / len2};
      double ll[3] = {len0, len1, len2};
      double bu[3] = {bu0, bu1, bu2};
      for (int e = 0; e < 3; ++e) {
        double src = -qp[e] + 0.0 * ll[e];
        double du = src / ll[e] - gB[e] - the[e] * gX[e];
        a.ou[ke3 + (int64_t)e * NC + gc] = a.trend_only ? du : bu[e] + a.h * du;
      }
    }
    m = nm;
    t = nt;
    u0 = nu0;
    u1 = nu1;
    u2 = nu2;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// trisk_stage_pair_kernel: two layer groups per block.  768 threads = two
// groups of 384; group 0 evaluates the even layers, group 1 the odd ones, in
// lock step, so each barrier serves two layers and a compute unit holds 12
// waves.  The launch-invariant geometry of the 19 x 19 region is staged once
// in LDS and shared by both groups; the Qperp weights of the owned cell (its
// own offsets 0-2 and those of E, NE, N for offsets 3-5) live in registers.
// Each thread has two roles: region cell lt (lt < 361) for staging, PV,
// fluxes and Bernoulli, and owned cell lt (lt < 256) for the trends, so only
// four waves per group run the Qperp / gradient phase.
// The hydrostatic column scan stays sequential per column: both groups
// advance the bottom pressure and geopotential through every layer, taking
// the other group's layer from its LDS planes, with the same operations in
// the same order as a single sweep.
// ---------------------------------------------------------------------------
constexpr int GROUP = 384;
constexpr int THREADS3 = 2 * GROUP;

struct GeoS {  // launch-invariant geometry of the staged region
  double len[3][RR], ped[3][RR], pedf[3][RR], area[6][RR], hinv[RR], cor[2][RR];
};
struct LayerS {  // one layer of one group; t holds t/m from phase 3 on
  double m[RR], t[RR], u[3][RR];
  double pvL[RR], pvU[RR];
  double qe[3][RR], hm[3][RR], ht[3][RR];
  double bern[RR], exn[RR];
};

__device__ __forceinline__ void weights_lds(const GeoS& gs, int y, int offs, double w[5]) {
  double acc0 = gs.area[offs][y];
  double acc1 = acc0 + gs.area[(1 + offs) % 6][y];
  double acc2 = acc1 + gs.area[(2 + offs) % 6][y];
  double acc3 = acc2 + gs.area[(3 + offs) % 6][y];
  double acc4 = acc3 + gs.area[(4 + offs) % 6][y];
  double hi = gs.hinv[y];
  w[0] = 0.5 - acc0 * hi;
  w[1] = -(0.5 - acc1 * hi);
  w[2] = 0.5 - acc2 * hi;
  w[3] = -(0.5 - acc3 * hi);
  w[4] = 0.5 - acc4 * hi;
}

// KEEP: also store the stage's fluxes, Bernoulli / Exner functions, velocity
// source and scalar trends (the k_* arrays) for the coarser levels of a
// hierarchy; SPLIT: the split-explicit stage, out = q + h (trend + theta1 g
// grad eta) on edges (u_star, barotropic_2d.f90:30-50, in split_axpy_kernel's
// order); DIFF: the diffusion hooks' values (phm, pht, phu) enter the fluxes
// and the velocity source.  Separate instantiations, so that the plain step
// carries none of them.
template <bool KEEP, bool SPLIT, bool DIFF>
__global__ __launch_bounds__(THREADS3) void trisk_stage_pair_kernel(StageArgs a_in) {
  const StageArgs a = member_args(a_in);
  __shared__ GeoS gs;
  __shared__ union {
    LayerS ls[2];
    double flat[2 * sizeof(LayerS) / sizeof(double)];
  } lsu;
  const int tid = threadIdx.x;
  const int grp = __builtin_amdgcn_readfirstlane(tid / GROUP);  // uniform over each wave
  const int lt = tid % GROUP;
  const int64_t NC = a.NC;
  const int S = a.S;
  const Geo& g = a.g;

  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int tile = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int tpl = a.ntile * a.ntile;
  const int lz = tile / tpl;
  const int rem = tile % tpl;
  const int i0 = (rem % a.ntile) * T, j0 = (rem / a.ntile) * T;
  const int64_t base = (int64_t)lz * S * S + (int64_t)(j0 - 1 + H) * S + (i0 - 1 + H);
  auto gcell = [&](int ri, int rj) -> int64_t { return base + (int64_t)rj * S + ri; };

  const bool cSW = (i0 == 0 && j0 == 0), cSE = (i0 + T == a.n && j0 == 0);
  const bool cNW = (i0 == 0 && j0 + T == a.n), cNE = (i0 + T == a.n && j0 + T == a.n);
  const bool corner = cSW || cSE || cNW || cNE;

  // region role
  const bool active = lt < RR;
  const int ri = lt % R, rj = lt / R;
  const int x = lt;
  const uint32_t xb = (uint32_t)(active ? gcell(ri, rj) : base) * 8u;  // byte offset of the cell in a plane
  const bool tri_cell = active && ri <= T + 1 && rj <= T + 1;
  const bool p_cell = active && ri >= 1 && rj >= 1 && ri <= T + 1 && rj <= T + 1;
  const bool pole = (lz == 0 && cSE && ri == T + 1 && rj == 1) || (lz == 5 && cNW && ri == 1 && rj == T + 1);
  // owned role
  const bool own = lt < T * T;
  const int oi = 1 + lt % T, oj = 1 + lt / T;
  const int y = oj * R + oi;
  const uint32_t ob = (uint32_t)gcell(oi, oj) * 8u;
  // The loads inside the layer loop are unconditional, at clamped addresses
  // (the values of idle threads and layers are dropped): with no branch
  // around them the compiler's wait counts stay exact and a prefetch is not
  // waited for before its use.
  const int lo = lt % (T * T);
  const uint32_t obl = (uint32_t)gcell(1 + lo % T, 1 + lo / T) * 8u;

  // --- launch-invariant geometry into LDS (each group stages half) ----------
  if (active) {
    if (grp == 0) {
      for (int e = 0; e < 3; ++e) {
        gs.len[e][x] = ldo(g.len + e * NC, xb);
        gs.ped[e][x] = ldo(g.ped + e * NC, xb);
        gs.pedf[e][x] = ldo(g.pedf + e * NC, xb);
      }
    } else {
      for (int p = 0; p < 6; ++p) gs.area[p][x] = ldo(g.area + p * NC, xb);
      gs.hinv[x] = ldo(g.hinv, xb);
      gs.cor[0][x] = ldo(g.cor, xb);
      gs.cor[1][x] = ldo(g.cor + NC, xb);
    }
  }

  // column state (cal_surf_press, set_surf_geopot): P = pressure at the
  // bottom of the next layer to scan, G = geopotential there.  The column
  // sum of (mass - temp) is taken in layer order from LDS, where each group
  // has stored (m - t) of every other layer (the layer planes are still free)
  double P = 0, G = 0;
  {
    constexpr int KMAX = (int)(sizeof(lsu.flat) / sizeof(double)) / RR;  // layers per round
    double s = 0.0;
    for (int k0 = 0; k0 < a.L; k0 += KMAX) {
      const int kn = min(KMAX, a.L - k0);
      if (active) {
        for (int kk = grp; kk < kn; kk += 2) {
          const int64_t ks = (int64_t)(k0 + kk) * NC;
          lsu.flat[kk * RR + x] = ldo(a.ym + ks, xb) - ldo(a.yt + ks, xb);
        }
      }
      __syncthreads();
      if (p_cell)
        for (int kk = 0; kk < kn; ++kk) s = s + lsu.flat[kk * RR + x];
      __syncthreads();
    }
    if (p_cell) {
      P = a.grav * s + a.p_top;
      G = a.grav * ldo(g.topo, xb);
    }
  }

  LayerS& sl = lsu.ls[grp];
  const LayerS& so = lsu.ls[1 - grp];
  // the group's first layer, staged here; later layers are staged at the end
  // of the previous iteration from registers loaded one iteration ahead
  double m, t, u0, u1, u2;
  {
    const int64_t ks = (int64_t)min(grp, a.L - 1) * NC, ke3 = 3 * ks;
    m = ldo(a.ym + ks, xb);
    t = ldo(a.yt + ks, xb);
    u0 = ldo(a.yu + ke3, xb);
    u1 = ldo(a.yu + ke3 + NC, xb);
    u2 = ldo(a.yu + ke3 + 2 * NC, xb);
  }
  if (active && grp < a.L) {
    sl.m[x] = m;
    sl.t[x] = t;
    sl.u[0][x] = u0;
    sl.u[1][x] = u1;
    sl.u[2][x] = u2;
  }
  __builtin_amdgcn_s_waitcnt(0);  // nothing outstanding enters the layer loop

  // Qperp weights of the neighbours' offsets 3-5 stay in registers; the
  // cell's own offsets 0-2 are rebuilt from LDS in each layer
  double w[6][5];
  if (own) {
    weights_lds(gs, y + 1, 3, w[3]);
    weights_lds(gs, y + R + 1, 4, w[4]);
    weights_lds(gs, y + R, 5, w[5]);
  }

  const int niter = (a.L + 1) / 2;
  for (int it = 0; it < niter; ++it) {
    const int k = 2 * it + grp;
    const bool lay = k < a.L;                 // uniform over the group
    const bool pair = 2 * it + 1 < a.L;       // both layers of the pair exist
    const int64_t ks = (int64_t)k * NC, ke3 = 3 * ks;
    __syncthreads();
    // --- 1. prefetch the group's next layer (its loads overlap phases 2-4) ----
    double nm, nt, nu0, nu1, nu2;
    {
      const int64_t ks2 = (int64_t)min(k + 2, a.L - 1) * NC, ke32 = 3 * ks2;
      nm = ldo(a.ym + ks2, xb);
      nt = ldo(a.yt + ks2, xb);
      nu0 = ldo(a.yu + ke32, xb);
      nu1 = ldo(a.yu + ke32 + NC, xb);
      nu2 = ldo(a.yu + ke32 + 2 * NC, xb);
    }
    // the layer's diffusion hooks (clamped layer, unconditional like the prefetch)
    double dpm[3] = {0.0, 0.0, 0.0}, dpt[3] = {0.0, 0.0, 0.0}, dpu[3] = {0.0, 0.0, 0.0};
    if (DIFF) {
      const int64_t kd3 = 3 * (int64_t)min(k, a.L - 1) * NC;
      for (int e = 0; e < 3; ++e) {
        if (a.phm) {
          dpm[e] = ldo(a.phm + kd3 + e * NC, xb);
          dpt[e] = ldo(a.pht + kd3 + e * NC, xb);
        }
        if (a.phu) dpu[e] = ldo(a.phu + kd3 + e * NC, obl);
      }
    }

    // --- 2. column scan; PV, fluxes; kinetic energy, Bernoulli, Exner ---------
    if (lay) {
      if (p_cell) {
        // layer 2it is group 0's; group 1 scans it from group 0's planes first
        double pr, phi_k;
        if (grp == 1) {
          const double mo = so.m[x], to = so.t[x];
          P = P - a.grav * (mo - to);
          G = G + a.grav * mo / (a.rho0 * 1.0);
        }
        {
          double p_upper = P - a.grav * (m - t);
          pr = interp(P, p_upper);
          double gpl = G;
          G = gpl + a.grav * m / (a.rho0 * 1.0);
          phi_k = interp(G, gpl);
          P = p_upper;
        }
        if (grp == 0 && pair) {
          const double mo = so.m[x], to = so.t[x];
          P = P - a.grav * (mo - to);
          G = G + a.grav * mo / (a.rho0 * 1.0);
        }
        const int xW = x - 1, xS = x - R, xSW = x - R - 1;
        double uS = sl.u[2][xS], uSW = sl.u[1][xSW], uW = sl.u[0][xW];
        double ke = ((u2 * gs.len[2][x]) * (u2 * gs.ped[2][x]) + (u1 * gs.len[1][x]) * (u1 * gs.ped[1][x]) +
                     (u0 * gs.len[0][x]) * (u0 * gs.ped[0][x]) + (uS * gs.len[2][xS]) * (uS * gs.ped[2][xS]) +
                     (uSW * gs.len[1][xSW]) * (uSW * gs.ped[1][xSW]) + (uW * gs.len[0][xW]) * (uW * gs.ped[0][xW])) *
                    gs.hinv[x] / 4.0;
        sl.bern[x] = ke + phi_k + pr / (a.rho0 * 1.0);
        sl.exn[x] = -phi_k;
        if (KEEP) {
          sto(a.k_bern + ks, xb, sl.bern[x]);
          sto(a.k_exn + ks, xb, sl.exn[x]);
        }
      }
      if (tri_cell) {
        const int xE = x + 1, xN = x + R, xNE = x + R + 1;
        double upRT = u0 * gs.len[0][x], upDG = u1 * gs.len[1][x], upUP = u2 * gs.len[2][x];
        double circL = upRT + sl.u[2][xE] * gs.len[2][xE] + upDG;
        double circU = -(upDG + upUP + sl.u[0][xN] * gs.len[0][xN]);
        double mE = sl.m[xE], mNE = sl.m[xNE], mN = sl.m[xN];
        double massL = m * gs.area[0][x] + mE * gs.area[2][xE] + mNE * gs.area[4][xNE];
        double massU = m * gs.area[1][x] + mNE * gs.area[3][xNE] + mN * gs.area[5][xN];
        sl.pvL[x] = (gs.cor[0][x] + circL) / massL;
        sl.pvU[x] = (gs.cor[1][x] + circU) / massU;
        const double f0 = gs.pedf[0][x], f1 = gs.pedf[1][x], f2 = gs.pedf[2][x];
        sl.hm[0][x] = u0 * f0 * interp(m, mE) + dpm[0];
        sl.hm[1][x] = u1 * f1 * interp(m, mNE) + dpm[1];
        sl.hm[2][x] = u2 * f2 * interp(m, mN) + dpm[2];
        sl.ht[0][x] = u0 * f0 * interp(t, sl.t[xE]) + dpt[0];
        sl.ht[1][x] = u1 * f1 * interp(t, sl.t[xNE]) + dpt[1];
        sl.ht[2][x] = u2 * f2 * interp(t, sl.t[xN]) + dpt[2];
        if (KEEP) {
          for (int e = 0; e < 3; ++e) {
            sto(a.k_hm + ke3 + e * NC, xb, sl.hm[e][x]);
            sto(a.k_ht + ke3 + e * NC, xb, sl.ht[e][x]);
          }
        }
      }
    }
    __syncthreads();

    // --- 3. base state of the update (q = y for a trend), consumed in phase 4;
    //        qe on edges; potential temperature t/m in place ------------------
    double bm, bt, bu0, bu1, bu2;
    {
      const int64_t ksb = (int64_t)min(k, a.L - 1) * NC, ke3b = 3 * ksb;
      bm = ldo(a.qm + ksb, obl);
      bt = ldo(a.qt + ksb, obl);
      bu0 = ldo(a.qu + ke3b, obl);
      bu1 = ldo(a.qu + ke3b + NC, obl);
      bu2 = ldo(a.qu + ke3b + 2 * NC, obl);
    }
    if (lay && active) {
      if (tri_cell) {
        if (rj >= 1) sl.qe[0][x] = interp(sl.pvU[x - R], sl.pvL[x]);
        sl.qe[1][x] = interp(sl.pvU[x], sl.pvL[x]);
        if (ri >= 1) sl.qe[2][x] = interp(sl.pvU[x], sl.pvL[x - 1]);
      }
      sl.t[x] = t / m;
    }
    __syncthreads();

    if (corner) {
      const bool mine[4] = {cSW, cSE, cNW, cNE};
      if (lay && lt < 4 && mine[lt])
        pentagon_fixup(lt, sl.u[0], sl.u[1], sl.u[2], sl.m, sl.qe[0], sl.qe[1], sl.qe[2], g, NC, base, S);
      __syncthreads();
    }

    // --- 4. trends of owned nodes and edges, pole master; RK update -----------
    if (lay && own) {
      const int xW = y - 1, xS = y - R, xSW = y - R - 1, xE = y + 1, xN = y + R, xNE = y + R + 1;
      const int xNW = y + R - 1, xSE = y - R + 1;
      const double hi = gs.hinv[y];
      double divm = (sl.hm[0][y] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][y] + sl.hm[2][y] - sl.hm[2][xS]) * hi;
      double divt = (sl.ht[0][y] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][y] + sl.ht[2][y] - sl.ht[2][xS]) * hi;
      double dmv = -divm, dtv = -divt;
      sto(a.om + ks, ob, a.trend_only ? dmv : bm + a.h * dmv);
      sto(a.ot + ks, ob, a.trend_only ? dtv : bt + a.h * dtv);
      if (KEEP) {
        sto(a.k_dm + ks, ob, dmv);
        sto(a.k_dt + ks, ob, dtv);
      }
      double qp[3];
      weights_lds(gs, y, 0, w[0]);
      weights_lds(gs, y, 1, w[1]);
      weights_lds(gs, y, 2, w[2]);
#define HM(e, xx) sl.hm[e][xx]
#define QE(e, xx) sl.qe[e][xx]
      double qr = QE(0, y);
      qp[0] = HM(1, y) * interp(QE(1, y), qr) * w[0][0] + HM(2, y) * interp(QE(2, y), qr) * w[0][1] +
              HM(0, xW) * interp(QE(0, xW), qr) * w[0][2] + HM(1, xSW) * interp(QE(1, xSW), qr) * w[0][3] +
              HM(2, xS) * interp(QE(2, xS), qr) * w[0][4] + HM(1, xS) * interp(QE(1, xS), qr) * w[3][0] +
              HM(2, xSE) * interp(QE(2, xSE), qr) * w[3][1] + HM(0, xE) * interp(QE(0, xE), qr) * w[3][2] +
              HM(1, xE) * interp(QE(1, xE), qr) * w[3][3] + HM(2, xE) * interp(QE(2, xE), qr) * w[3][4];
      double qd = QE(1, y);
      qp[1] = HM(2, y) * interp(QE(2, y), qd) * w[1][0] + HM(0, xW) * interp(QE(0, xW), qd) * w[1][1] +
              HM(1, xSW) * interp(QE(1, xSW), qd) * w[1][2] + HM(2, xS) * interp(QE(2, xS), qd) * w[1][3] +
              HM(0, y) * interp(QE(0, y), qd) * w[1][4] + HM(2, xE) * interp(QE(2, xE), qd) * w[4][0] +
              HM(0, xNE) * interp(QE(0, xNE), qd) * w[4][1] + HM(1, xNE) * interp(QE(1, xNE), qd) * w[4][2] +
              HM(2, xNE) * interp(QE(2, xNE), qd) * w[4][3] + HM(0, xN) * interp(QE(0, xN), qd) * w[4][4];
      double qu = QE(2, y);
      qp[2] = HM(0, xW) * interp(QE(0, xW), qu) * w[2][0] + HM(1, xSW) * interp(QE(1, xSW), qu) * w[2][1] +
              HM(2, xS) * interp(QE(2, xS), qu) * w[2][2] + HM(0, y) * interp(QE(0, y), qu) * w[2][3] +
              HM(1, y) * interp(QE(1, y), qu) * w[2][4] + HM(0, xN) * interp(QE(0, xN), qu) * w[5][0] +
              HM(1, xN) * interp(QE(1, xN), qu) * w[5][1] + HM(2, xN) * interp(QE(2, xN), qu) * w[5][2] +
              HM(0, xNW) * interp(QE(0, xNW), qu) * w[5][3] + HM(1, xW) * interp(QE(1, xW), qu) * w[5][4];
#undef HM
#undef QE
      const double th0 = sl.t[y];
      double the[3] = {interp(th0, sl.t[xE]), interp(th0, sl.t[xNE]), interp(th0, sl.t[xN])};
      const double b0 = sl.bern[y], x0 = sl.exn[y];
      const double ll[3] = {gs.len[0][y], gs.len[1][y], gs.len[2][y]};
      double gB[3] = {(sl.bern[xE] - b0) / ll[0], (b0 - sl.bern[xNE]) / ll[1], (sl.bern[xN] - b0) / ll[2]};
      double gX[3] = {(sl.exn[xE] - x0) / ll[0], (x0 - sl.exn[xNE]) / ll[1], (sl.exn[xN] - x0) / ll[2]};
      const double bu[3] = {bu0, bu1, bu2};
      double ge[3] = {0.0, 0.0, 0.0};
      if (SPLIT) {  // grad_eta_e: g (eta(E) - eta) / len, g (eta - eta(NE)) / len, g (eta(N) - eta) / len
        const double e0 = ldo(a.eta, ob);
        ge[0] = a.grav * (ldo(a.eta, ob + 8u) - e0) / ll[0];
        ge[1] = a.grav * (e0 - ldo(a.eta, ob + 8u * (uint32_t)(S + 1))) / ll[1];
        ge[2] = a.grav * (ldo(a.eta, ob + 8u * (uint32_t)S) - e0) / ll[2];
      }
      for (int e = 0; e < 3; ++e) {
        double src = -qp[e] + dpu[e] * ll[e];
        double du = src / ll[e] - gB[e] - the[e] * gX[e];
        sto(a.ou + ke3 + e * NC, ob, a.trend_only ? du : bu[e] + a.h * (SPLIT ? du + a.th1 * ge[e] : du));
        if (KEEP) sto(a.k_src + ke3 + e * NC, ob, src);
      }
    }
    if (!(lay && own)) {
      // threads that store nothing still wait for their base-state loads
      // here, so no later phase inherits an unknown wait on them
      asm volatile("" ::"v"(bm), "v"(bt), "v"(bu0), "v"(bu1), "v"(bu2));
    }
    if (lay && pole) {
      const double hi = gs.hinv[x];
      const int xW = x - 1, xS = x - R, xSW = x - R - 1;
      double divm = (sl.hm[0][x] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][x] + sl.hm[2][x] - sl.hm[2][xS]) * hi;
      double divt = (sl.ht[0][x] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][x] + sl.ht[2][x] - sl.ht[2][xS]) * hi;
      double dmv = -divm, dtv = -divt;
      sto(a.om + ks, xb, a.trend_only ? dmv : ldo(a.qm + ks, xb) + a.h * dmv);
      sto(a.ot + ks, xb, a.trend_only ? dtv : ldo(a.qt + ks, xb) + a.h * dtv);
      if (KEEP) {
        sto(a.k_dm + ks, xb, dmv);
        sto(a.k_dt + ks, xb, dtv);
      }
    }
    __syncthreads();
    // --- 5. stage the group's next layer --------------------------------------
    if (active && k + 2 < a.L) {
      sl.m[x] = nm;
      sl.t[x] = nt;
      sl.u[0][x] = nu0;
      sl.u[1][x] = nu1;
      sl.u[2][x] = nu2;
    }
    m = nm;
    t = nt;
    u0 = nu0;
    u1 = nu1;
    u2 = nu2;
  }
}

// halo := owner (nodes: mass and temp; edges: velocity with sign), all
// layers; F64 state slots and the FP32 stage buffers (F = float).  Thread =
// (list entry, chunk of GF_KC layers), entries fastest, so neighbouring
// threads copy neighbouring list entries of one layer plane; a thread reads
// its chunk's sources before it writes (a list's destinations are never its
// sources, the entries are independent).
constexpr int GF_KC = 8;
template <class F>
__global__ void ghost_fill_kernel(int L, int64_t NC, const int64_t* __restrict__ nd, int nn,
                                  const int64_t* __restrict__ ed, const int8_t* __restrict__ es, int ne, F* m, F* t,
                                  F* u) {
  const int64_t E = (int64_t)nn + ne;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= E * ((L + GF_KC - 1) / GF_KC)) return;
  const int64_t e = idx % E;
  const int k0 = (int)(idx / E) * GF_KC, nk = min(GF_KC, L - k0);
  if (e < nn) {
    const int64_t dst = nd[2 * e], src = nd[2 * e + 1];
    F vm[GF_KC], vt[GF_KC];
#pragma unroll
    for (int i = 0; i < GF_KC; ++i)
      if (i < nk) {
        vm[i] = m[(int64_t)(k0 + i) * NC + src];
        vt[i] = t[(int64_t)(k0 + i) * NC + src];
      }
#pragma unroll
    for (int i = 0; i < GF_KC; ++i)
      if (i < nk) {
        m[(int64_t)(k0 + i) * NC + dst] = vm[i];
        t[(int64_t)(k0 + i) * NC + dst] = vt[i];
      }
  } else {
    const int64_t j = e - nn;
    const int64_t dst = ed[2 * j], src = ed[2 * j + 1];  // plane indices e*NC + cell
    const F sg = (F)es[j];
    F vu[GF_KC];
#pragma unroll
    for (int i = 0; i < GF_KC; ++i)
      if (i < nk) vu[i] = u[(int64_t)(k0 + i) * 3 * NC + src];
#pragma unroll
    for (int i = 0; i < GF_KC; ++i)
      if (i < nk) u[(int64_t)(k0 + i) * 3 * NC + dst] = sg * vu[i];
  }
}

// Restrict_velo (wavelet.f90:1154-1174) over the owned coarse nodes: a
// coarse edge is the mean of the two fine edges that bisect it, in the
// lozenge frame of the child (2i, 2j) -- RT: child, child + E; DG:
// child + NE, child; UP: child, child + N.  One thread per (node, layer).
__global__ void restrict_velo_kernel(int nc, int Sc, int Sf, int64_t NCc, int64_t NCf, int L,
                                     const double* __restrict__ uf, double* __restrict__ uc) {
  const int64_t per = (int64_t)10 * nc * nc;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * L) return;
  const int k = (int)(t / per);
  const int64_t r = t % per;
  const int lz = (int)(r / ((int64_t)nc * nc)), rr = (int)(r % ((int64_t)nc * nc));
  const int j = rr / nc, i = rr % nc;
  const int64_t p = (int64_t)lz * Sc * Sc + (int64_t)(j + 2) * Sc + (i + 2);
  const int64_t c = (int64_t)lz * Sf * Sf + (int64_t)(2 * j + 2) * Sf + (2 * i + 2);
  const double* f = uf + (int64_t)k * 3 * NCf;
  double* o = uc + (int64_t)k * 3 * NCc;
  o[p] = 0.5 * (f[c] + f[c + 1]);
  o[NCc + p] = 0.5 * (f[NCf + c + Sf + 1] + f[NCf + c]);
  o[2 * NCc + p] = 0.5 * (f[2 * NCf + c] + f[2 * NCf + c + Sf]);
}

// copy frozen halo slots (never ghost filled) from one state to another
__global__ void copy_frozen_kernel(int L, int64_t NC, const int64_t* __restrict__ fn, int nn,
                                   const int64_t* __restrict__ fe, int ne, const double* sm_, const double* st_,
                                   const double* su_, double* dm_, double* dt_, double* du_) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)(nn + ne) * L;
  if (idx >= total) return;
  int k = (int)(idx % L);
  int64_t e = idx / L;
  if (e < nn) {
    int64_t c = fn[e];
    dm_[(int64_t)k * NC + c] = sm_[(int64_t)k * NC + c];
    dt_[(int64_t)k * NC + c] = st_[(int64_t)k * NC + c];
  } else {
    int64_t p = fe[e - nn];
    du_[(int64_t)k * 3 * NC + p] = su_[(int64_t)k * 3 * NC + p];
  }
}


// ---------------------------------------------------------------------------
// FP32-bulk stage kernel (precision mode WTG_PRECISION_FP32_BULK)
// ---------------------------------------------------------------------------
// The same tile, layer pairing and phases as trisk_stage_pair_kernel, with
// the bulk of the stencils in FP32: the velocity, mass and temperature planes
// of the flux stencils, circulation, potential vorticity, qe, mass and
// temperature fluxes, kinetic energy, divergence and the Qperp sum.  FP64
// stays where cancellation or accumulation would cost accuracy: the state
// itself, the hydrostatic column scan (pressure and geopotential), Bernoulli
// and Exner and their edge differences, and the update of the state.
// The intermediate RK stages are kept as FP32 increments over the FP64 state
// q: stage k evaluates the trend on y = q + d_k (d_1 = 0) and writes
// d_{k+1} = h_{k+1} * trend(y) in FP32, the last stage writes
// q + h * trend(y) in FP64.  An increment is small beside the state, so its
// FP32 rounding is far below that of an FP32 copy of the state, and nothing
// rounded accumulates in the prognostic variables.
struct GeoF {          // FP32 copies of the stencil geometry
  const float* len;    // [3][NC]
  const float* lp;     // [3][NC] len * pedlen (kinetic energy)
  const float* pedf;   // [3][NC]
  const float* area;   // [6][NC]
  const float* hinv;   // [NC]
  const float* cor;    // [2][NC]
};

template <class To>
struct MixedArgs {
  int n, S, L, ntile;
  int64_t NC;
  const double *qm, *qt, *qu;  // FP64 state (ghost filled)
  const float *dm, *dt, *du;   // FP32 stage increment over it (ghost filled; unused in the first stage)
  To *om, *ot, *ou;            // increment of the next stage (To = float) or the new state (To = double)
  double h;
  GeoF f;
  Geo g;                       // FP64 geometry: topography, pentagon corrections
  double grav, inv_rho0, p_top;
};

template <class To>
__device__ __forceinline__ MixedArgs<To> member_args(const MixedArgs<To>& a0) {
  MixedArgs<To> a = a0;
  const int64_t s1 = (int64_t)blockIdx.y * a0.L * a0.NC, s3 = 3 * s1;
  a.qm += s1; a.qt += s1; a.qu += s3;
  a.dm += s1; a.dt += s1; a.du += s3;
  a.om += s1; a.ot += s1; a.ou += s3;
  a.g.topo += (int64_t)blockIdx.y * a0.NC;
  return a;
}

struct GeoSF {
  float len[3][RR], lp[3][RR], area[6][RR], hinv[RR];
};
struct LayerSF {
  double bern[RR], exn[RR];
  double mt[RR], md[RR];  // m - t and m in FP64 for the other group's column scan
  float m[RR], t[RR], u[3][RR];
  float th[RR];           // potential temperature t/m
  float pvL[RR], pvU[RR];
  float qe[3][RR], hm[3][RR], ht[3][RR];
};

// load / store of element `cell` of a plane (32-bit offset from a uniform base)
template <class F>
__device__ __forceinline__ F ldc(const F* p, uint32_t cell) {
  return *reinterpret_cast<const F*>(reinterpret_cast<const char*>(p) + cell * (uint32_t)sizeof(F));
}
template <class F>
__device__ __forceinline__ void stc(F* p, uint32_t cell, F v) {
  *reinterpret_cast<F*>(reinterpret_cast<char*>(p) + cell * (uint32_t)sizeof(F)) = v;
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fint(float a, float b) { return 0.5f * (a + b); }

__device__ __forceinline__ void weights_f32(const GeoSF& gs, int y, int offs, float w[5]) {
  float acc0 = gs.area[offs][y];
  float acc1 = acc0 + gs.area[(1 + offs) % 6][y];
  float acc2 = acc1 + gs.area[(2 + offs) % 6][y];
  float acc3 = acc2 + gs.area[(3 + offs) % 6][y];
  float acc4 = acc3 + gs.area[(4 + offs) % 6][y];
  float hi = gs.hinv[y];
  w[0] = 0.5f - acc0 * hi;
  w[1] = -(0.5f - acc1 * hi);
  w[2] = 0.5f - acc2 * hi;
  w[3] = -(0.5f - acc3 * hi);
  w[4] = 0.5f - acc4 * hi;
}

// value of the evaluation state y = q + d at plane offset `ks`, cell `c`
template <bool HasD>
__device__ __forceinline__ double yval(const double* q, const float* d, int64_t ks, uint32_t c) {
  return HasD ? ldc(q + ks, c) + (double)ldc(d + ks, c) : ldc(q + ks, c);
}

// Velocity trend of edge e of the owned cell at LDS index yy (Qperp, the
// Bernoulli and Exner gradients), in FP32 from the layer planes; the Qperp
// weights of the cell (offset e) and of its neighbour (offset e + 3) are built
// from the area planes
__device__ __forceinline__ float edge_du(const LayerSF& sl, const GeoSF& gs, int yy, int e) {
  const int xW = yy - 1, xS = yy - R, xSW = yy - R - 1, xE = yy + 1, xN = yy + R, xNE = yy + R + 1;
  const int xNW = yy + R - 1, xSE = yy - R + 1;
  const float th0 = sl.th[yy];
  const double b0 = sl.bern[yy], x0 = sl.exn[yy];
  float w0[5], w1[5], qp, the, dB, dX, il;
#define HM(e, xx) sl.hm[e][xx]
#define QE(e, xx) sl.qe[e][xx]
  if (e == 0) {
    weights_f32(gs, yy, 0, w0);
    weights_f32(gs, yy + 1, 3, w1);
    const float qr = QE(0, yy);
    qp = HM(1, yy) * fint(QE(1, yy), qr) * w0[0] + HM(2, yy) * fint(QE(2, yy), qr) * w0[1] +
         HM(0, xW) * fint(QE(0, xW), qr) * w0[2] + HM(1, xSW) * fint(QE(1, xSW), qr) * w0[3] +
         HM(2, xS) * fint(QE(2, xS), qr) * w0[4] + HM(1, xS) * fint(QE(1, xS), qr) * w1[0] +
         HM(2, xSE) * fint(QE(2, xSE), qr) * w1[1] + HM(0, xE) * fint(QE(0, xE), qr) * w1[2] +
         HM(1, xE) * fint(QE(1, xE), qr) * w1[3] + HM(2, xE) * fint(QE(2, xE), qr) * w1[4];
    the = fint(th0, sl.th[xE]);
    dB = (float)(sl.bern[xE] - b0);
    dX = (float)(sl.exn[xE] - x0);
    il = frcp(gs.len[0][yy]);
  } else if (e == 1) {
    weights_f32(gs, yy, 1, w0);
    weights_f32(gs, yy + R + 1, 4, w1);
    const float qd = QE(1, yy);
    qp = HM(2, yy) * fint(QE(2, yy), qd) * w0[0] + HM(0, xW) * fint(QE(0, xW), qd) * w0[1] +
         HM(1, xSW) * fint(QE(1, xSW), qd) * w0[2] + HM(2, xS) * fint(QE(2, xS), qd) * w0[3] +
         HM(0, yy) * fint(QE(0, yy), qd) * w0[4] + HM(2, xE) * fint(QE(2, xE), qd) * w1[0] +
         HM(0, xNE) * fint(QE(0, xNE), qd) * w1[1] + HM(1, xNE) * fint(QE(1, xNE), qd) * w1[2] +
         HM(2, xNE) * fint(QE(2, xNE), qd) * w1[3] + HM(0, xN) * fint(QE(0, xN), qd) * w1[4];
    the = fint(th0, sl.th[xNE]);
    dB = (float)(b0 - sl.bern[xNE]);
    dX = (float)(x0 - sl.exn[xNE]);
    il = frcp(gs.len[1][yy]);
  } else {
    weights_f32(gs, yy, 2, w0);
    weights_f32(gs, yy + R, 5, w1);
    const float qu = QE(2, yy);
    qp = HM(0, xW) * fint(QE(0, xW), qu) * w0[0] + HM(1, xSW) * fint(QE(1, xSW), qu) * w0[1] +
         HM(2, xS) * fint(QE(2, xS), qu) * w0[2] + HM(0, yy) * fint(QE(0, yy), qu) * w0[3] +
         HM(1, yy) * fint(QE(1, yy), qu) * w0[4] + HM(0, xN) * fint(QE(0, xN), qu) * w1[0] +
         HM(1, xN) * fint(QE(1, xN), qu) * w1[1] + HM(2, xN) * fint(QE(2, xN), qu) * w1[2] +
         HM(0, xNW) * fint(QE(0, xNW), qu) * w1[3] + HM(1, xW) * fint(QE(1, xW), qu) * w1[4];
    the = fint(th0, sl.th[xN]);
    dB = (float)(sl.bern[xN] - b0);
    dX = (float)(sl.exn[xN] - x0);
    il = frcp(gs.len[2][yy]);
  }
#undef HM
#undef QE
  return (-qp - dB - the * dX) * il;
}

// HasD: the stage has an increment (stages 2-4); To: float = the next
// stage's increment, double = the new state; Edge: phase 4 as edge tasks
// (each group's 384 threads take the 3 x 256 edge trends, two each, and the
// first 256 also the node trends) instead of one owned cell per thread
template <bool HasD, class To, bool Edge>
__global__ __launch_bounds__(THREADS3) void trisk_stage_mixed_kernel(MixedArgs<To> a_in) {
#pragma clang fp contract(fast)
  const MixedArgs<To> a = member_args(a_in);
  constexpr bool kState = sizeof(To) == 8;
  __shared__ GeoSF gs;
  __shared__ union {
    LayerSF ls[2];
    double flat[2 * sizeof(LayerSF) / sizeof(double)];
  } lsu;
  const int tid = threadIdx.x;
  const int grp = __builtin_amdgcn_readfirstlane(tid / GROUP);
  const int lt = tid % GROUP;
  const int64_t NC = a.NC;
  const int S = a.S;

  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int tile = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int tpl = a.ntile * a.ntile;
  const int lz = tile / tpl;
  const int rem = tile % tpl;
  const int i0 = (rem % a.ntile) * T, j0 = (rem / a.ntile) * T;
  const int64_t base = (int64_t)lz * S * S + (int64_t)(j0 - 1 + H) * S + (i0 - 1 + H);
  auto gcell = [&](int ri, int rj) -> int64_t { return base + (int64_t)rj * S + ri; };

  const bool cSW = (i0 == 0 && j0 == 0), cSE = (i0 + T == a.n && j0 == 0);
  const bool cNW = (i0 == 0 && j0 + T == a.n), cNE = (i0 + T == a.n && j0 + T == a.n);
  const bool corner = cSW || cSE || cNW || cNE;

  const bool active = lt < RR;
  const int ri = lt % R, rj = lt / R;
  const int x = lt;
  const uint32_t xc = (uint32_t)(active ? gcell(ri, rj) : base);
  const bool tri_cell = active && ri <= T + 1 && rj <= T + 1;
  const bool p_cell = active && ri >= 1 && rj >= 1 && ri <= T + 1 && rj <= T + 1;
  const bool pole = (lz == 0 && cSE && ri == T + 1 && rj == 1) || (lz == 5 && cNW && ri == 1 && rj == T + 1);
  const bool own = lt < T * T;
  const int oi = 1 + lt % T, oj = 1 + lt / T;
  const int y = oj * R + oi;
  const uint32_t oc = (uint32_t)gcell(oi, oj);
  const int lo = lt % (T * T);
  const uint32_t ocl = (uint32_t)gcell(1 + lo % T, 1 + lo / T);

  // --- geometry: planes in LDS, the flux lengths and Coriolis in registers ---
  float pf0 = 0.f, pf1 = 0.f, pf2 = 0.f, cr0 = 0.f, cr1 = 0.f;
  if (active) {
    if (grp == 0) {
      for (int e = 0; e < 3; ++e) {
        gs.len[e][x] = ldc(a.f.len + e * NC, xc);
        gs.lp[e][x] = ldc(a.f.lp + e * NC, xc);
      }
      gs.hinv[x] = ldc(a.f.hinv, xc);
    } else {
      for (int p = 0; p < 6; ++p) gs.area[p][x] = ldc(a.f.area + p * NC, xc);
    }
    pf0 = ldc(a.f.pedf, xc);
    pf1 = ldc(a.f.pedf + NC, xc);
    pf2 = ldc(a.f.pedf + 2 * NC, xc);
    cr0 = ldc(a.f.cor, xc);
    cr1 = ldc(a.f.cor + NC, xc);
  }

  // column state in FP64 (cal_surf_press, set_surf_geopot): P = pressure at
  // the bottom of the next layer, G = geopotential there
  double P = 0, G = 0;
  {
    constexpr int KMAX = (int)(sizeof(lsu.flat) / sizeof(double)) / RR;
    double s = 0.0;
    for (int k0 = 0; k0 < a.L; k0 += KMAX) {
      const int kn = min(KMAX, a.L - k0);
      if (active) {
        for (int kk = grp; kk < kn; kk += 2) {
          const int64_t ks = (int64_t)(k0 + kk) * NC;
          lsu.flat[kk * RR + x] = yval<HasD>(a.qm, a.dm, ks, xc) - yval<HasD>(a.qt, a.dt, ks, xc);
        }
      }
      __syncthreads();
      if (p_cell)
        for (int kk = 0; kk < kn; ++kk) s = s + lsu.flat[kk * RR + x];
      __syncthreads();
    }
    if (p_cell) {
      P = a.grav * s + a.p_top;
      G = a.grav * ldc(a.g.topo, xc);
    }
  }
  const double g_rho = a.grav * a.inv_rho0;

  LayerSF& sl = lsu.ls[grp];
  const LayerSF& so = lsu.ls[1 - grp];
  double m, t;  // the group's layer: mass and temp in FP64, velocity in FP32
  float u0, u1, u2;
  {
    const int64_t ks = (int64_t)min(grp, a.L - 1) * NC, ke3 = 3 * ks;
    m = yval<HasD>(a.qm, a.dm, ks, xc);
    t = yval<HasD>(a.qt, a.dt, ks, xc);
    u0 = (float)yval<HasD>(a.qu, a.du, ke3, xc);
    u1 = (float)yval<HasD>(a.qu, a.du, ke3 + NC, xc);
    u2 = (float)yval<HasD>(a.qu, a.du, ke3 + 2 * NC, xc);
  }
  if (active && grp < a.L) {
    sl.mt[x] = m - t;
    sl.md[x] = m;
    sl.m[x] = (float)m;
    sl.t[x] = (float)t;
    sl.u[0][x] = u0;
    sl.u[1][x] = u1;
    sl.u[2][x] = u2;
  }
  __syncthreads();

  float w[6][5];
  if (!Edge && own) {
    weights_f32(gs, y + 1, 3, w[3]);
    weights_f32(gs, y + R + 1, 4, w[4]);
    weights_f32(gs, y + R, 5, w[5]);
  }
  // edge tasks of this thread (Edge): task t -> edge t / 256 of owned cell t % 256
  const int tA = lt, tB = lt + GROUP;
  const int eA = tA / (T * T), eB = tB / (T * T);
  const int cA = tA % (T * T), cB = tB % (T * T);
  const int yA = (1 + cA / T) * R + 1 + cA % T, yB = (1 + cB / T) * R + 1 + cB % T;
  const uint32_t gA = (uint32_t)gcell(1 + cA % T, 1 + cA / T), gB = (uint32_t)gcell(1 + cB % T, 1 + cB / T);

  const int niter = (a.L + 1) / 2;
  for (int it = 0; it < niter; ++it) {
    const int k = 2 * it + grp;
    const bool lay = k < a.L;
    const bool pair = 2 * it + 1 < a.L;
    const int64_t ks = (int64_t)k * NC, ke3 = 3 * ks;
    __syncthreads();
    // --- 1. prefetch the group's next layer (q and d stay apart until phase 5,
    //        so nothing waits for these loads before then) ---------------------
    double nm, nt, nu0, nu1, nu2;
    float dm = 0.f, dt = 0.f, du0 = 0.f, du1 = 0.f, du2 = 0.f;
    {
      const int64_t ks2 = (int64_t)min(k + 2, a.L - 1) * NC, ke32 = 3 * ks2;
      nm = ldc(a.qm + ks2, xc);
      nt = ldc(a.qt + ks2, xc);
      nu0 = ldc(a.qu + ke32, xc);
      nu1 = ldc(a.qu + ke32 + NC, xc);
      nu2 = ldc(a.qu + ke32 + 2 * NC, xc);
      if (HasD) {
        dm = ldc(a.dm + ks2, xc);
        dt = ldc(a.dt + ks2, xc);
        du0 = ldc(a.du + ke32, xc);
        du1 = ldc(a.du + ke32 + NC, xc);
        du2 = ldc(a.du + ke32 + 2 * NC, xc);
      }
    }

    // --- 2. column scan (FP64); PV, fluxes, kinetic energy (FP32) -------------
    const float mf = (float)m, tf = (float)t;
    if (lay) {
      if (p_cell) {
        double pr, phi_k;
        if (grp == 1) {
          P = P - a.grav * so.mt[x];
          G = G + g_rho * so.md[x];
        }
        {
          double p_upper = P - a.grav * (m - t);
          pr = 0.5 * (P + p_upper);
          double gpl = G;
          G = gpl + g_rho * m;
          phi_k = 0.5 * (G + gpl);
          P = p_upper;
        }
        if (grp == 0 && pair) {
          P = P - a.grav * so.mt[x];
          G = G + g_rho * so.md[x];
        }
        const int xW = x - 1, xS = x - R, xSW = x - R - 1;
        const float uS = sl.u[2][xS], uSW = sl.u[1][xSW], uW = sl.u[0][xW];
        const float ke = (u2 * u2 * gs.lp[2][x] + u1 * u1 * gs.lp[1][x] + u0 * u0 * gs.lp[0][x] +
                          uS * uS * gs.lp[2][xS] + uSW * uSW * gs.lp[1][xSW] + uW * uW * gs.lp[0][xW]) *
                         (0.25f * gs.hinv[x]);
        sl.bern[x] = (double)ke + phi_k + pr * a.inv_rho0;
        sl.exn[x] = -phi_k;
      }
      if (tri_cell) {
        const int xE = x + 1, xN = x + R, xNE = x + R + 1;
        const float upRT = u0 * gs.len[0][x], upDG = u1 * gs.len[1][x], upUP = u2 * gs.len[2][x];
        const float circL = upRT + sl.u[2][xE] * gs.len[2][xE] + upDG;
        const float circU = -(upDG + upUP + sl.u[0][xN] * gs.len[0][xN]);
        const float mE = sl.m[xE], mNE = sl.m[xNE], mN = sl.m[xN];
        const float massL = mf * gs.area[0][x] + mE * gs.area[2][xE] + mNE * gs.area[4][xNE];
        const float massU = mf * gs.area[1][x] + mNE * gs.area[3][xNE] + mN * gs.area[5][xN];
        sl.pvL[x] = (cr0 + circL) * frcp(massL);
        sl.pvU[x] = (cr1 + circU) * frcp(massU);
        const float f0 = u0 * pf0, f1 = u1 * pf1, f2 = u2 * pf2;
        sl.hm[0][x] = f0 * fint(mf, mE);
        sl.hm[1][x] = f1 * fint(mf, mNE);
        sl.hm[2][x] = f2 * fint(mf, mN);
        sl.ht[0][x] = f0 * fint(tf, sl.t[xE]);
        sl.ht[1][x] = f1 * fint(tf, sl.t[xNE]);
        sl.ht[2][x] = f2 * fint(tf, sl.t[xN]);
      }
    }
    __syncthreads();

    // --- 3. FP64 base state of the update (last stage); qe; t/m in place ------
    double bm = 0, bt = 0, bu0 = 0, bu1 = 0, bu2 = 0;
    if (kState) {
      const int64_t ksb = (int64_t)min(k, a.L - 1) * NC, ke3b = 3 * ksb;
      bm = ldc(a.qm + ksb, ocl);
      bt = ldc(a.qt + ksb, ocl);
      if (Edge) {
        bu0 = ldc(a.qu + ke3b + eA * NC, gA);
        bu1 = ldc(a.qu + ke3b + eB * NC, gB);
      } else {
        bu0 = ldc(a.qu + ke3b, ocl);
        bu1 = ldc(a.qu + ke3b + NC, ocl);
        bu2 = ldc(a.qu + ke3b + 2 * NC, ocl);
      }
    }
    if (lay && active) {
      if (tri_cell) {
        if (rj >= 1) sl.qe[0][x] = fint(sl.pvU[x - R], sl.pvL[x]);
        sl.qe[1][x] = fint(sl.pvU[x], sl.pvL[x]);
        if (ri >= 1) sl.qe[2][x] = fint(sl.pvU[x], sl.pvL[x - 1]);
      }
      sl.th[x] = (float)(t / m);
    }
    __syncthreads();

    if (corner) {
      const bool mine[4] = {cSW, cSE, cNW, cNE};
      if (lay && lt < 4 && mine[lt])
        pentagon_fixup(lt, sl.u[0], sl.u[1], sl.u[2], sl.m, sl.qe[0], sl.qe[1], sl.qe[2], a.g, NC, base, S);
      __syncthreads();
    }

    // --- 4. trends of owned nodes and edges, pole master; update -------------
    if (Edge && lay) {
      if (own) {
        const int xW = y - 1, xS = y - R, xSW = y - R - 1;
        const float hi = gs.hinv[y];
        const float divm =
            (sl.hm[0][y] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][y] + sl.hm[2][y] - sl.hm[2][xS]) * hi;
        const float divt =
            (sl.ht[0][y] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][y] + sl.ht[2][y] - sl.ht[2][xS]) * hi;
        stc(a.om + ks, oc, (To)(bm - a.h * (double)divm));
        stc(a.ot + ks, oc, (To)(bt - a.h * (double)divt));
      }
      const float duA = edge_du(sl, gs, yA, eA);
      stc(a.ou + ke3 + eA * NC, gA, (To)(bu0 + a.h * (double)duA));
      const float duB = edge_du(sl, gs, yB, eB);
      stc(a.ou + ke3 + eB * NC, gB, (To)(bu1 + a.h * (double)duB));
    }
    if (!Edge && lay && own) {
      const int xW = y - 1, xS = y - R, xSW = y - R - 1, xE = y + 1, xN = y + R, xNE = y + R + 1;
      const int xNW = y + R - 1, xSE = y - R + 1;
      const float hi = gs.hinv[y];
      const float divm =
          (sl.hm[0][y] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][y] + sl.hm[2][y] - sl.hm[2][xS]) * hi;
      const float divt =
          (sl.ht[0][y] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][y] + sl.ht[2][y] - sl.ht[2][xS]) * hi;
      stc(a.om + ks, oc, (To)(bm - a.h * (double)divm));
      stc(a.ot + ks, oc, (To)(bt - a.h * (double)divt));
      float qp[3];
      weights_f32(gs, y, 0, w[0]);
      weights_f32(gs, y, 1, w[1]);
      weights_f32(gs, y, 2, w[2]);
#define HM(e, xx) sl.hm[e][xx]
#define QE(e, xx) sl.qe[e][xx]
      const float qr = QE(0, y);
      qp[0] = HM(1, y) * fint(QE(1, y), qr) * w[0][0] + HM(2, y) * fint(QE(2, y), qr) * w[0][1] +
              HM(0, xW) * fint(QE(0, xW), qr) * w[0][2] + HM(1, xSW) * fint(QE(1, xSW), qr) * w[0][3] +
              HM(2, xS) * fint(QE(2, xS), qr) * w[0][4] + HM(1, xS) * fint(QE(1, xS), qr) * w[3][0] +
              HM(2, xSE) * fint(QE(2, xSE), qr) * w[3][1] + HM(0, xE) * fint(QE(0, xE), qr) * w[3][2] +
              HM(1, xE) * fint(QE(1, xE), qr) * w[3][3] + HM(2, xE) * fint(QE(2, xE), qr) * w[3][4];
      const float qd = QE(1, y);
      qp[1] = HM(2, y) * fint(QE(2, y), qd) * w[1][0] + HM(0, xW) * fint(QE(0, xW), qd) * w[1][1] +
              HM(1, xSW) * fint(QE(1, xSW), qd) * w[1][2] + HM(2, xS) * fint(QE(2, xS), qd) * w[1][3] +
              HM(0, y) * fint(QE(0, y), qd) * w[1][4] + HM(2, xE) * fint(QE(2, xE), qd) * w[4][0] +
              HM(0, xNE) * fint(QE(0, xNE), qd) * w[4][1] + HM(1, xNE) * fint(QE(1, xNE), qd) * w[4][2] +
              HM(2, xNE) * fint(QE(2, xNE), qd) * w[4][3] + HM(0, xN) * fint(QE(0, xN), qd) * w[4][4];
      const float qu = QE(2, y);
      qp[2] = HM(0, xW) * fint(QE(0, xW), qu) * w[2][0] + HM(1, xSW) * fint(QE(1, xSW), qu) * w[2][1] +
              HM(2, xS) * fint(QE(2, xS), qu) * w[2][2] + HM(0, y) * fint(QE(0, y), qu) * w[2][3] +
              HM(1, y) * fint(QE(1, y), qu) * w[2][4] + HM(0, xN) * fint(QE(0, xN), qu) * w[5][0] +
              HM(1, xN) * fint(QE(1, xN), qu) * w[5][1] + HM(2, xN) * fint(QE(2, xN), qu) * w[5][2] +
              HM(0, xNW) * fint(QE(0, xNW), qu) * w[5][3] + HM(1, xW) * fint(QE(1, xW), qu) * w[5][4];
#undef HM
#undef QE
      const float th0 = sl.th[y];
      const float the[3] = {fint(th0, sl.th[xE]), fint(th0, sl.th[xNE]), fint(th0, sl.th[xN])};
      const double b0 = sl.bern[y], x0 = sl.exn[y];
      const float il[3] = {frcp(gs.len[0][y]), frcp(gs.len[1][y]), frcp(gs.len[2][y])};
      const float dB[3] = {(float)(sl.bern[xE] - b0), (float)(b0 - sl.bern[xNE]), (float)(sl.bern[xN] - b0)};
      const float dX[3] = {(float)(sl.exn[xE] - x0), (float)(x0 - sl.exn[xNE]), (float)(sl.exn[xN] - x0)};
      const double bu[3] = {bu0, bu1, bu2};
      for (int e = 0; e < 3; ++e) {
        const float du = (-qp[e] - dB[e] - the[e] * dX[e]) * il[e];
        stc(a.ou + ke3 + e * NC, oc, (To)(bu[e] + a.h * (double)du));
      }
    }
    if (kState && !(lay && (own || Edge))) asm volatile("" ::"v"(bm), "v"(bt), "v"(bu0), "v"(bu1), "v"(bu2));
    if (kState && Edge && lay && !own) asm volatile("" ::"v"(bm), "v"(bt));
    if (lay && pole) {
      const float hi = gs.hinv[x];
      const int xW = x - 1, xS = x - R, xSW = x - R - 1;
      const float divm =
          (sl.hm[0][x] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][x] + sl.hm[2][x] - sl.hm[2][xS]) * hi;
      const float divt =
          (sl.ht[0][x] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][x] + sl.ht[2][x] - sl.ht[2][xS]) * hi;
      const double pm = kState ? ldc(a.qm + ks, xc) : 0.0, pt = kState ? ldc(a.qt + ks, xc) : 0.0;
      stc(a.om + ks, xc, (To)(pm - a.h * (double)divm));
      stc(a.ot + ks, xc, (To)(pt - a.h * (double)divt));
    }
    // --- 5. stage the group's next layer (phase 4 reads none of these planes,
    //        and the next iteration's barrier orders them before phase 2) ------
    m = nm + (double)dm;
    t = nt + (double)dt;
    u0 = (float)(nu0 + (double)du0);
    u1 = (float)(nu1 + (double)du1);
    u2 = (float)(nu2 + (double)du2);
    if (active && k + 2 < a.L) {
      sl.mt[x] = m - t;
      sl.md[x] = m;
      sl.m[x] = (float)m;
      sl.t[x] = (float)t;
      sl.u[0][x] = u0;
      sl.u[1][x] = u1;
      sl.u[2][x] = u2;
    }
  }
}

// One layer at a time per block (the FP32-bulk kernel's second shape): a
// single group of 384 threads per tile walks the layers bottom to top, so the
// column scan stays in registers and LDS holds one layer (49 KB): three
// blocks share a CU and cover each other's barriers, where the layer-pair
// shape holds one 768-thread block per CU in lock step.
struct LayerS1 {
  double bern[RR], exn[RR];
  float m[RR], t[RR], u[3][RR], th[RR];
  float pvL[RR], pvU[RR];
  float qe[3][RR], hm[3][RR], ht[3][RR];
};
constexpr int THREADS1 = 384;

template <bool HasD, class To>
__global__ __launch_bounds__(THREADS1, 5) void trisk_stage_mixed1_kernel(MixedArgs<To> a_in) {
#pragma clang fp contract(fast)
  const MixedArgs<To> a = member_args(a_in);
  constexpr bool kState = sizeof(To) == 8;
  __shared__ GeoSF gs;
  __shared__ union {
    LayerS1 l;
    double flat[sizeof(LayerS1) / sizeof(double)];
  } lsu;
  LayerS1& sl = lsu.l;
  const int lt = threadIdx.x;
  const int64_t NC = a.NC;
  const int S = a.S;

  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int tile = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const int tpl = a.ntile * a.ntile;
  const int lz = tile / tpl;
  const int rem = tile % tpl;
  const int i0 = (rem % a.ntile) * T, j0 = (rem / a.ntile) * T;
  const int64_t base = (int64_t)lz * S * S + (int64_t)(j0 - 1 + H) * S + (i0 - 1 + H);
  auto gcell = [&](int ri, int rj) -> int64_t { return base + (int64_t)rj * S + ri; };

  const bool cSW = (i0 == 0 && j0 == 0), cSE = (i0 + T == a.n && j0 == 0);
  const bool cNW = (i0 == 0 && j0 + T == a.n), cNE = (i0 + T == a.n && j0 + T == a.n);
  const bool corner = cSW || cSE || cNW || cNE;

  const bool active = lt < RR;
  const int ri = lt % R, rj = lt / R;
  const int x = lt;
  const uint32_t xc = (uint32_t)(active ? gcell(ri, rj) : base);
  const bool tri_cell = active && ri <= T + 1 && rj <= T + 1;
  const bool p_cell = active && ri >= 1 && rj >= 1 && ri <= T + 1 && rj <= T + 1;
  const bool pole = (lz == 0 && cSE && ri == T + 1 && rj == 1) || (lz == 5 && cNW && ri == 1 && rj == T + 1);
  const bool own = lt < T * T;
  const int oi = 1 + lt % T, oj = 1 + lt / T;
  const int y = oj * R + oi;
  const uint32_t oc = (uint32_t)gcell(oi, oj);
  const int lo = lt % (T * T);
  const uint32_t ocl = (uint32_t)gcell(1 + lo % T, 1 + lo / T);

  float pf0 = 0.f, pf1 = 0.f, pf2 = 0.f, cr0 = 0.f, cr1 = 0.f;
  if (active) {
    for (int e = 0; e < 3; ++e) {
      gs.len[e][x] = ldc(a.f.len + e * NC, xc);
      gs.lp[e][x] = ldc(a.f.lp + e * NC, xc);
    }
    for (int p = 0; p < 6; ++p) gs.area[p][x] = ldc(a.f.area + p * NC, xc);
    gs.hinv[x] = ldc(a.f.hinv, xc);
    pf0 = ldc(a.f.pedf, xc);
    pf1 = ldc(a.f.pedf + NC, xc);
    pf2 = ldc(a.f.pedf + 2 * NC, xc);
    cr0 = ldc(a.f.cor, xc);
    cr1 = ldc(a.f.cor + NC, xc);
  }

  // column state in FP64 (cal_surf_press, set_surf_geopot)
  double P = 0, G = 0;
  {
    constexpr int KMAX = (int)(sizeof(lsu.flat) / sizeof(double)) / RR;
    double s = 0.0;
    for (int k0 = 0; k0 < a.L; k0 += KMAX) {
      const int kn = min(KMAX, a.L - k0);
      if (active) {
        for (int kk = 0; kk < kn; ++kk) {
          const int64_t ks = (int64_t)(k0 + kk) * NC;
          lsu.flat[kk * RR + x] = yval<HasD>(a.qm, a.dm, ks, xc) - yval<HasD>(a.qt, a.dt, ks, xc);
        }
      }
      __syncthreads();
      if (p_cell)
        for (int kk = 0; kk < kn; ++kk) s = s + lsu.flat[kk * RR + x];
      __syncthreads();
    }
    if (p_cell) {
      P = a.grav * s + a.p_top;
      G = a.grav * ldc(a.g.topo, xc);
    }
  }
  const double g_rho = a.grav * a.inv_rho0;

  double m, t;
  float u0, u1, u2;
  m = yval<HasD>(a.qm, a.dm, 0, xc);
  t = yval<HasD>(a.qt, a.dt, 0, xc);
  u0 = (float)yval<HasD>(a.qu, a.du, 0, xc);
  u1 = (float)yval<HasD>(a.qu, a.du, NC, xc);
  u2 = (float)yval<HasD>(a.qu, a.du, 2 * NC, xc);

  // the Qperp weights are rebuilt from LDS in each layer (registers are
  // what limits this shape to three blocks per CU)
  float w[6][5];

  for (int k = 0; k < a.L; ++k) {
    const int64_t ks = (int64_t)k * NC, ke3 = 3 * ks;
    // --- 1. stage the layer; prefetch the next one and the update's base ------
    const float mf = (float)m, tf = (float)t;
    if (active) {
      sl.m[x] = mf;
      sl.t[x] = tf;
      sl.u[0][x] = u0;
      sl.u[1][x] = u1;
      sl.u[2][x] = u2;
    }
    double nm, nt, nu0, nu1, nu2;
    float dm = 0.f, dt = 0.f, du0 = 0.f, du1 = 0.f, du2 = 0.f;
    {
      const int64_t ks2 = (int64_t)min(k + 1, a.L - 1) * NC, ke32 = 3 * ks2;
      nm = ldc(a.qm + ks2, xc);
      nt = ldc(a.qt + ks2, xc);
      nu0 = ldc(a.qu + ke32, xc);
      nu1 = ldc(a.qu + ke32 + NC, xc);
      nu2 = ldc(a.qu + ke32 + 2 * NC, xc);
      if (HasD) {
        dm = ldc(a.dm + ks2, xc);
        dt = ldc(a.dt + ks2, xc);
        du0 = ldc(a.du + ke32, xc);
        du1 = ldc(a.du + ke32 + NC, xc);
        du2 = ldc(a.du + ke32 + 2 * NC, xc);
      }
    }
    double bm = 0, bt = 0, bu0 = 0, bu1 = 0, bu2 = 0;
    if (kState) {
      bm = ldc(a.qm + ks, ocl);
      bt = ldc(a.qt + ks, ocl);
      bu0 = ldc(a.qu + ke3, ocl);
      bu1 = ldc(a.qu + ke3 + NC, ocl);
      bu2 = ldc(a.qu + ke3 + 2 * NC, ocl);
    }
    double pr = 0, phi_k = 0;
    if (p_cell) {
      double p_upper = P - a.grav * (m - t);
      pr = 0.5 * (P + p_upper);
      double gpl = G;
      G = gpl + g_rho * m;
      phi_k = 0.5 * (G + gpl);
      P = p_upper;
    }
    __syncthreads();

    // --- 2. PV, fluxes; kinetic energy, Bernoulli, Exner ----------------------
    if (p_cell) {
      const int xW = x - 1, xS = x - R, xSW = x - R - 1;
      const float uS = sl.u[2][xS], uSW = sl.u[1][xSW], uW = sl.u[0][xW];
      const float ke = (u2 * u2 * gs.lp[2][x] + u1 * u1 * gs.lp[1][x] + u0 * u0 * gs.lp[0][x] +
                        uS * uS * gs.lp[2][xS] + uSW * uSW * gs.lp[1][xSW] + uW * uW * gs.lp[0][xW]) *
                       (0.25f * gs.hinv[x]);
      sl.bern[x] = (double)ke + phi_k + pr * a.inv_rho0;
      sl.exn[x] = -phi_k;
    }
    if (tri_cell) {
      const int xE = x + 1, xN = x + R, xNE = x + R + 1;
      const float upRT = u0 * gs.len[0][x], upDG = u1 * gs.len[1][x], upUP = u2 * gs.len[2][x];
      const float circL = upRT + sl.u[2][xE] * gs.len[2][xE] + upDG;
      const float circU = -(upDG + upUP + sl.u[0][xN] * gs.len[0][xN]);
      const float mE = sl.m[xE], mNE = sl.m[xNE], mN = sl.m[xN];
      const float massL = mf * gs.area[0][x] + mE * gs.area[2][xE] + mNE * gs.area[4][xNE];
      const float massU = mf * gs.area[1][x] + mNE * gs.area[3][xNE] + mN * gs.area[5][xN];
      sl.pvL[x] = (cr0 + circL) * frcp(massL);
      sl.pvU[x] = (cr1 + circU) * frcp(massU);
      const float f0 = u0 * pf0, f1 = u1 * pf1, f2 = u2 * pf2;
      sl.hm[0][x] = f0 * fint(mf, mE);
      sl.hm[1][x] = f1 * fint(mf, mNE);
      sl.hm[2][x] = f2 * fint(mf, mN);
      sl.ht[0][x] = f0 * fint(tf, sl.t[xE]);
      sl.ht[1][x] = f1 * fint(tf, sl.t[xNE]);
      sl.ht[2][x] = f2 * fint(tf, sl.t[xN]);
    }
    if (active) sl.th[x] = (float)(t / m);
    __syncthreads();

    // --- 3. qe ------------------------------------------------------------------
    if (tri_cell) {
      if (rj >= 1) sl.qe[0][x] = fint(sl.pvU[x - R], sl.pvL[x]);
      sl.qe[1][x] = fint(sl.pvU[x], sl.pvL[x]);
      if (ri >= 1) sl.qe[2][x] = fint(sl.pvU[x], sl.pvL[x - 1]);
    }
    __syncthreads();
    if (corner) {
      const bool mine[4] = {cSW, cSE, cNW, cNE};
      if (lt < 4 && mine[lt])
        pentagon_fixup(lt, sl.u[0], sl.u[1], sl.u[2], sl.m, sl.qe[0], sl.qe[1], sl.qe[2], a.g, NC, base, S);
      __syncthreads();
    }

    // --- 4. trends of owned nodes and edges, pole master; update --------------
    // (the next layer's staging writes only m, t, u, which this phase does not
    // read; its barrier orders this phase before the next phase 2)
    if (own) {
      const int xW = y - 1, xS = y - R, xSW = y - R - 1, xE = y + 1, xN = y + R, xNE = y + R + 1;
      const int xNW = y + R - 1, xSE = y - R + 1;
      const float hi = gs.hinv[y];
      const float divm =
          (sl.hm[0][y] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][y] + sl.hm[2][y] - sl.hm[2][xS]) * hi;
      const float divt =
          (sl.ht[0][y] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][y] + sl.ht[2][y] - sl.ht[2][xS]) * hi;
      stc(a.om + ks, oc, (To)(bm - a.h * (double)divm));
      stc(a.ot + ks, oc, (To)(bt - a.h * (double)divt));
      float qp[3];
      weights_f32(gs, y, 0, w[0]);
      weights_f32(gs, y, 1, w[1]);
      weights_f32(gs, y, 2, w[2]);
      weights_f32(gs, y + 1, 3, w[3]);
      weights_f32(gs, y + R + 1, 4, w[4]);
      weights_f32(gs, y + R, 5, w[5]);
#define HM(e, xx) sl.hm[e][xx]
#define QE(e, xx) sl.qe[e][xx]
      const float qr = QE(0, y);
      qp[0] = HM(1, y) * fint(QE(1, y), qr) * w[0][0] + HM(2, y) * fint(QE(2, y), qr) * w[0][1] +
              HM(0, xW) * fint(QE(0, xW), qr) * w[0][2] + HM(1, xSW) * fint(QE(1, xSW), qr) * w[0][3] +
              HM(2, xS) * fint(QE(2, xS), qr) * w[0][4] + HM(1, xS) * fint(QE(1, xS), qr) * w[3][0] +
              HM(2, xSE) * fint(QE(2, xSE), qr) * w[3][1] + HM(0, xE) * fint(QE(0, xE), qr) * w[3][2] +
              HM(1, xE) * fint(QE(1, xE), qr) * w[3][3] + HM(2, xE) * fint(QE(2, xE), qr) * w[3][4];
      const float qd = QE(1, y);
      qp[1] = HM(2, y) * fint(QE(2, y), qd) * w[1][0] + HM(0, xW) * fint(QE(0, xW), qd) * w[1][1] +
              HM(1, xSW) * fint(QE(1, xSW), qd) * w[1][2] + HM(2, xS) * fint(QE(2, xS), qd) * w[1][3] +
              HM(0, y) * fint(QE(0, y), qd) * w[1][4] + HM(2, xE) * fint(QE(2, xE), qd) * w[4][0] +
              HM(0, xNE) * fint(QE(0, xNE), qd) * w[4][1] + HM(1, xNE) * fint(QE(1, xNE), qd) * w[4][2] +
              HM(2, xNE) * fint(QE(2, xNE), qd) * w[4][3] + HM(0, xN) * fint(QE(0, xN), qd) * w[4][4];
      const float qu = QE(2, y);
      qp[2] = HM(0, xW) * fint(QE(0, xW), qu) * w[2][0] + HM(1, xSW) * fint(QE(1, xSW), qu) * w[2][1] +
              HM(2, xS) * fint(QE(2, xS), qu) * w[2][2] + HM(0, y) * fint(QE(0, y), qu) * w[2][3] +
              HM(1, y) * fint(QE(1, y), qu) * w[2][4] + HM(0, xN) * fint(QE(0, xN), qu) * w[5][0] +
              HM(1, xN) * fint(QE(1, xN), qu) * w[5][1] + HM(2, xN) * fint(QE(2, xN), qu) * w[5][2] +
              HM(0, xNW) * fint(QE(0, xNW), qu) * w[5][3] + HM(1, xW) * fint(QE(1, xW), qu) * w[5][4];
#undef HM
#undef QE
      const float th0 = sl.th[y];
      const float the[3] = {fint(th0, sl.th[xE]), fint(th0, sl.th[xNE]), fint(th0, sl.th[xN])};
      const double b0 = sl.bern[y], x0 = sl.exn[y];
      const float il[3] = {frcp(gs.len[0][y]), frcp(gs.len[1][y]), frcp(gs.len[2][y])};
      const float dB[3] = {(float)(sl.bern[xE] - b0), (float)(b0 - sl.bern[xNE]), (float)(sl.bern[xN] - b0)};
      const float dX[3] = {(float)(sl.exn[xE] - x0), (float)(x0 - sl.exn[xNE]), (float)(sl.exn[xN] - x0)};
      const double bu[3] = {bu0, bu1, bu2};
      for (int e = 0; e < 3; ++e) {
        const float du = (-qp[e] - dB[e] - the[e] * dX[e]) * il[e];
        stc(a.ou + ke3 + e * NC, oc, (To)(bu[e] + a.h * (double)du));
      }
    }
    if (kState && !own) asm volatile("" ::"v"(bm), "v"(bt), "v"(bu0), "v"(bu1), "v"(bu2));
    if (pole) {
      const float hi = gs.hinv[x];
      const int xW = x - 1, xS = x - R, xSW = x - R - 1;
      const float divm =
          (sl.hm[0][x] - sl.hm[0][xW] + sl.hm[1][xSW] - sl.hm[1][x] + sl.hm[2][x] - sl.hm[2][xS]) * hi;
      const float divt =
          (sl.ht[0][x] - sl.ht[0][xW] + sl.ht[1][xSW] - sl.ht[1][x] + sl.ht[2][x] - sl.ht[2][xS]) * hi;
      const double pm = kState ? ldc(a.qm + ks, xc) : 0.0, pt = kState ? ldc(a.qt + ks, xc) : 0.0;
      stc(a.om + ks, xc, (To)(pm - a.h * (double)divm));
      stc(a.ot + ks, xc, (To)(pt - a.h * (double)divt));
    }
    m = nm + (double)dm;
    t = nt + (double)dt;
    u0 = (float)(nu0 + (double)du0);
    u1 = (float)(nu1 + (double)du1);
    u2 = (float)(nu2 + (double)du2);
  }
}

// FP32 copies of the stencil geometry for the FP32-bulk kernel
__global__ void geo_f32_kernel(int64_t NC, const double* __restrict__ len, const double* __restrict__ ped,
                               const double* __restrict__ pedf, const double* __restrict__ area,
                               const double* __restrict__ hinv, const double* __restrict__ cor, float* flen,
                               float* flp, float* fpedf, float* farea, float* fhinv, float* fcor) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= NC) return;
  for (int e = 0; e < 3; ++e) {
    flen[e * NC + c] = (float)len[e * NC + c];
    flp[e * NC + c] = (float)(len[e * NC + c] * ped[e * NC + c]);
    fpedf[e * NC + c] = (float)pedf[e * NC + c];
  }
  for (int p = 0; p < 6; ++p) farea[p * NC + c] = (float)area[p * NC + c];
  fhinv[c] = (float)hinv[c];
  fcor[c] = (float)cor[c];
  fcor[NC + c] = (float)cor[NC + c];
}

// ---------------------------------------------------------------------------
// split-explicit step with implicit free surface (mode split)
// ---------------------------------------------------------------------------
// barotropic_correction (barotropic_2d.f90:52-116) in place over the masters
// (the first nmaster entries of the node compute list, dst == src): layer
// dilation of mass / temp to the free surface eta; porosity 1, mean state 0
__global__ void baro_correction_kernel(int nmaster, const int64_t* __restrict__ ncl, int L, int64_t NC, double rho0,
                                       const double* __restrict__ eta, const double* __restrict__ topo, double* m,
                                       double* t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmaster) return;
  const int64_t c = ncl[2 * i + 1];
  double h = 0.0;  // total_height (cal_height: rho_dz / porous_density)
  for (int k = 0; k < L; ++k) h = h + (0.0 + m[(int64_t)k * NC + c]) / (rho0 * 1.0);
  const double e = eta[c] / 1.0, b = topo[c];
  for (int k = 0; k < L; ++k) {
    const int64_t x = (int64_t)k * NC + c;
    double rho_dz = 0.0 + m[x];
    const double theta = (0.0 + t[x]) / rho_dz;
    const double mm = (e - b) / h * rho_dz - 0.0;
    m[x] = mm;
    rho_dz = 0.0 + mm;
    t[x] = rho_dz * theta - 0.0;
  }
}

// g grad(eta) on edge e of cell c (grad_eta, step1 itype 3, ops.f90:217-220)
__device__ __forceinline__ double grad_eta_e(const double* __restrict__ eta, const double* __restrict__ len,
                                             int64_t NC, int S, int64_t c, int e, double grav) {
  if (e == 0) return grav * (eta[c + 1] / 1.0 - eta[c] / 1.0) / len[c];
  if (e == 1) return grav * (eta[c] / 1.0 - eta[c + S + 1] / 1.0) / len[NC + c];
  return grav * (eta[c + S] / 1.0 - eta[c] / 1.0) / len[2 * NC + c];
}

// scalar_star and u_star (barotropic_2d.f90:9-50) over the masters:
// out = base + h tr (scalars), out = base + h (tr + theta1 g grad eta) (edges);
// the trend is FP64, or FP32 in the FP32-bulk mode (the sums stay FP64)
template <class Tr>
__global__ void split_axpy_kernel(int nmaster, const int64_t* __restrict__ ncl, int L, int64_t NC, int S, double h,
                                  double th1, double grav, const double* __restrict__ eta,
                                  const double* __restrict__ len, const double* bm, const double* bt,
                                  const double* bu, const Tr* tm, const Tr* tt, const Tr* tu, double* om,
                                  double* ot, double* ou) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nmaster * L) return;
  const int k = (int)(idx / nmaster), i = (int)(idx % nmaster);
  const int64_t c = ncl[2 * i + 1], x = (int64_t)k * NC + c;
  om[x] = bm[x] + h * (double)tm[x];
  ot[x] = bt[x] + h * (double)tt[x];
  for (int e = 0; e < 3; ++e) {
    const int64_t y = ((int64_t)k * 3 + e) * NC + c;
    ou[y] = bu[y] + h * ((double)tu[y] + th1 * grad_eta_e(eta, len, NC, S, c, e, grav));
  }
}

// u_update (barotropic_2d.f90:131-149) in place over the masters: u -= theta1 dt g grad eta
__global__ void u_update_kernel(int nmaster, const int64_t* __restrict__ ncl, int L, int64_t NC, int S, double th1dt,
                                double grav, const double* __restrict__ eta, const double* __restrict__ len,
                                double* u) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nmaster * L) return;
  const int k = (int)(idx / nmaster), i = (int)(idx % nmaster);
  const int64_t c = ncl[2 * i + 1];
  for (int e = 0; e < 3; ++e) {
    const int64_t y = ((int64_t)k * 3 + e) * NC + c;
    u[y] = u[y] - th1dt * grad_eta_e(eta, len, NC, S, c, e, grav);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host runtime
// ---------------------------------------------------------------------------
thread_local std::string g_create_error;

int wtg_fail(wtg_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_create_error = buf;
  return code;
}

namespace {

int fail(wtg_ctx* c, int code, const char* msg) { return wtg_fail(c, code, "%s", msg); }
template <class... A>
int fail(wtg_ctx* c, int code, const char* fmt, A... args) {
  return wtg_fail(c, code, fmt, args...);
}
#define HIPCHK WTG_HIPCHK
template <class Tp>
int dalloc(wtg_ctx* c, Tp** p, size_t count) {
  return wtg_dalloc(c, p, count);
}

int check_slot(wtg_ctx* c, int slot) {
  if (!c) return WTG_E_ARG;
  if (slot < 0 || slot >= c->nslots) return fail(c, WTG_E_ARG, "slot %d out of range [0,%d)", slot, c->nslots);
  return 0;
}

template <class F>
int launch_ghost_fill(wtg_ctx* c, F* m, F* t, F* u) {
  const int L = c->params.nlayers * c->nmembers;
  int64_t total = (int64_t)(c->nn + c->ne) * ((L + GF_KC - 1) / GF_KC);
  if (total == 0) return 0;
  int threads = 256;
  int64_t blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(ghost_fill_kernel<F>, dim3((unsigned)blocks), dim3(threads), 0, c->stream, L, c->info.ncells,
                     c->d_nd, c->nn, c->d_ed, c->d_es, c->ne, m, t, u);
  HIPCHK(c, hipGetLastError());
  return 0;
}
int launch_ghost_fill(wtg_ctx* c, const wtg_ctx::Slot& s) { return launch_ghost_fill(c, s.m, s.t, s.u); }

int launch_copy_frozen(wtg_ctx* c, const wtg_ctx::Slot& from, const wtg_ctx::Slot& to) {
  const int L = c->params.nlayers * c->nmembers;
  int64_t total = (int64_t)(c->nfn + c->nfe) * L;
  if (total == 0) return 0;
  int threads = 256;
  int64_t blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(copy_frozen_kernel, dim3((unsigned)blocks), dim3(threads), 0, c->stream, L, c->info.ncells,
                     c->d_fn, c->nfn, c->d_fe, c->nfe, from.m, from.t, from.u, to.m, to.t, to.u);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_stage(wtg_ctx* c, const wtg_ctx::Slot& y, const wtg_ctx::Slot& q, const wtg_ctx::Slot& out, double h,
                 int trend_only, const WtgStageKeep* keep = nullptr, const double* split_eta = nullptr) {
  StageArgs a{};
  if (split_eta) {
    if (c->stage_version != 3 || keep || trend_only)
      return fail(c, WTG_E_STATE, "the fused split-explicit stage needs the layer-pair kernel");
    a.eta = split_eta;
    a.th1 = c->fs_theta1;
  }
  if (c->dif_on) {
    if (c->stage_version != 3 || keep)
      return fail(c, WTG_E_STATE, "diffusion needs the layer-pair stage kernel on a single level");
    if (int rc = wtg_diffusion_prepass(c, y.m, y.t, y.u)) return rc;
    a.phm = c->dif_phm;
    a.pht = c->dif_pht;
    a.phu = c->dif_phu;
  }
  if (keep) {
    if (c->stage_version != 3 || c->nmembers != 1)
      return fail(c, WTG_E_STATE, "kept stage quantities need the layer-pair kernel on a single-member context");
    a.k_hm = keep->hm;
    a.k_ht = keep->ht;
    a.k_bern = keep->bern;
    a.k_exn = keep->exn;
    a.k_src = keep->src;
    a.k_dm = keep->dm;
    a.k_dt = keep->dt;
  }
  a.n = c->info.n;
  a.S = c->info.S;
  a.L = c->params.nlayers;
  a.ntile = a.n / T;
  a.NC = c->info.ncells;
  a.spole = c->info.south_pole;
  a.npole = c->info.north_pole;
  a.ym = y.m;
  a.yt = y.t;
  a.yu = y.u;
  a.qm = q.m;
  a.qt = q.t;
  a.qu = q.u;
  a.om = out.m;
  a.ot = out.t;
  a.ou = out.u;
  a.h = h;
  a.trend_only = trend_only;
  a.g = Geo{c->d_len, c->d_ped, c->d_pedf, c->d_area, c->d_hinv, c->d_tri, c->d_cor, c->d_topo};
  a.grav = c->params.grav;
  a.rho0 = c->params.ref_density;
  a.p_top = c->params.p_top;
  unsigned blocks = (unsigned)(10 * a.ntile * a.ntile);
  const dim3 grid(blocks, (unsigned)c->nmembers);
  if (c->stage_version == 1)
    hipLaunchKernelGGL(trisk_stage_kernel, grid, dim3(THREADS), 0, c->stream, a);
  else if (c->stage_version == 2)
    hipLaunchKernelGGL(trisk_stage_kernel2, grid, dim3(THREADS2), 0, c->stream, a);
  else if (keep)
    hipLaunchKernelGGL((trisk_stage_pair_kernel<true, false, false>), grid, dim3(THREADS3), 0, c->stream, a);
  else if (split_eta && c->dif_on)
    hipLaunchKernelGGL((trisk_stage_pair_kernel<false, true, true>), grid, dim3(THREADS3), 0, c->stream, a);
  else if (split_eta)
    hipLaunchKernelGGL((trisk_stage_pair_kernel<false, true, false>), grid, dim3(THREADS3), 0, c->stream, a);
  else if (c->dif_on)
    hipLaunchKernelGGL((trisk_stage_pair_kernel<false, false, true>), grid, dim3(THREADS3), 0, c->stream, a);
  else
    hipLaunchKernelGGL((trisk_stage_pair_kernel<false, false, false>), grid, dim3(THREADS3), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// FP32 geometry copies and the two FP32 stage buffers of the FP32-bulk mode
int f32_ready(wtg_ctx* c) {
  if (c->f32_done) return 0;
  const int64_t NC = c->info.ncells;
  const size_t L = (size_t)c->params.nlayers * c->nmembers;
  int rc = 0;
  if ((rc = dalloc(c, &c->f_len, 3 * NC)) || (rc = dalloc(c, &c->f_lp, 3 * NC)) ||
      (rc = dalloc(c, &c->f_pedf, 3 * NC)) || (rc = dalloc(c, &c->f_area, 6 * NC)) ||
      (rc = dalloc(c, &c->f_hinv, NC)) || (rc = dalloc(c, &c->f_cor, 2 * NC)))
    return rc;
  for (auto& s : c->fstage) {
    if ((rc = dalloc(c, &s.m, L * NC)) || (rc = dalloc(c, &s.t, L * NC)) || (rc = dalloc(c, &s.u, 3 * L * NC)))
      return rc;
    HIPCHK(c, hipMemsetAsync(s.m, 0, L * NC * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(s.t, 0, L * NC * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(s.u, 0, 3 * L * NC * sizeof(float), c->stream));
  }
  hipLaunchKernelGGL(geo_f32_kernel, dim3((unsigned)((NC + 255) / 256)), dim3(256), 0, c->stream, NC, c->d_len,
                     c->d_ped, c->d_pedf, c->d_area, c->d_hinv, c->d_cor, c->f_len, c->f_lp, c->f_pedf, c->f_area,
                     c->f_hinv, c->f_cor);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->f32_done = true;
  return 0;
}

template <bool HasD, class To>
int launch_stage_mixed(wtg_ctx* c, const wtg_ctx::Slot& q, const wtg_ctx::SlotF* d, To* om, To* ot, To* ou,
                       double h) {
  MixedArgs<To> a{};
  a.n = c->info.n;
  a.S = c->info.S;
  a.L = c->params.nlayers;
  a.ntile = a.n / T;
  a.NC = c->info.ncells;
  a.qm = q.m;
  a.qt = q.t;
  a.qu = q.u;
  // the first stage has no increment: its pointers are never read
  a.dm = d ? d->m : c->fstage[0].m;
  a.dt = d ? d->t : c->fstage[0].t;
  a.du = d ? d->u : c->fstage[0].u;
  a.om = om;
  a.ot = ot;
  a.ou = ou;
  a.h = h;
  a.f = GeoF{c->f_len, c->f_lp, c->f_pedf, c->f_area, c->f_hinv, c->f_cor};
  a.g = Geo{c->d_len, c->d_ped, c->d_pedf, c->d_area, c->d_hinv, c->d_tri, c->d_cor, c->d_topo};
  a.grav = c->params.grav;
  a.inv_rho0 = 1.0 / c->params.ref_density;
  a.p_top = c->params.p_top;
  const dim3 grid((unsigned)(10 * a.ntile * a.ntile), (unsigned)c->nmembers);
  if (c->mixed_shape == WTG_SHAPE_ONE_LAYER)
    hipLaunchKernelGGL((trisk_stage_mixed1_kernel<HasD, To>), grid, dim3(THREADS1), 0, c->stream, a);
  else if (c->mixed_shape == WTG_SHAPE_EDGE_TASKS)
    hipLaunchKernelGGL((trisk_stage_mixed_kernel<HasD, To, true>), grid, dim3(THREADS3), 0, c->stream, a);
  else
    hipLaunchKernelGGL((trisk_stage_mixed_kernel<HasD, To, false>), grid, dim3(THREADS3), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// The autotuner of the FP32-bulk stage: times every launch shape on the
// context's own state (an intermediate stage, 3 launches each after a warm-up
// launch) and keeps the fastest.  The launches write only the second FP32
// increment buffer, which the next RK stage overwrites before reading it.
int tune_mixed_shape(wtg_ctx* c, const wtg_ctx::Slot& q) {
  const int shapes[3] = {WTG_SHAPE_LAYER_PAIRS, WTG_SHAPE_EDGE_TASKS, WTG_SHAPE_ONE_LAYER};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return fail(c, WTG_E_DEVICE, "hipEventCreate failed");
  }
  int rc = 0, best = shapes[0];
  float best_ms = 0.f;
  auto& f1 = c->fstage[0];
  auto& f2 = c->fstage[1];
  for (int i = 0; i < 3 && !rc; ++i) {
    c->mixed_shape = shapes[i];
    rc = launch_stage_mixed<true, float>(c, q, &f1, f2.m, f2.t, f2.u, 1.0);
    if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(c, WTG_E_DEVICE, "event record failed");
    for (int r = 0; r < 3 && !rc; ++r) rc = launch_stage_mixed<true, float>(c, q, &f1, f2.m, f2.t, f2.u, 1.0);
    if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
      rc = fail(c, WTG_E_DEVICE, "event synchronise failed");
    float ms = 0.f;
    if (!rc) hipEventElapsedTime(&ms, e0, e1);
    c->tuned_ms[i] = ms / 3.0f;
    if (!rc && (i == 0 || ms < best_ms)) {
      best_ms = ms;
      best = shapes[i];
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->mixed_shape = best;
  c->mixed_tuned = true;
  return rc;
}

// FP32-bulk RK4: the low-storage stages with FP32 increments d over the FP64
// state q (y_k = q + d_k, d_{k+1} = h_{k+1} f(y_k)); the last stage writes
// q + dt f(y_4) into the other FP64 buffer, because every stage reads q at
// the halo of its tile, where the neighbouring tiles would be updating it in
// place.  The state ping-pongs between the slot and the context's first FP64
// stage buffer and is copied back after an odd number of steps.  Halo slots
// that are never ghost filled keep d = 0 (the buffers start zeroed and
// stages write owned cells and pole masters only).
int rk4_mixed_steps(wtg_ctx* c, int slot, double dt, int nsteps, hipEvent_t* ev) {
  const wtg_ctx::Slot& sol = c->slots[slot];
  const wtg_ctx::Slot* cur = &sol;
  const wtg_ctx::Slot* nxt = &c->slots[c->nslots];
  auto &f1 = c->fstage[0], &f2 = c->fstage[1];
  int launch = 0;
  auto rec = [&](int k) -> int {
    if (ev) HIPCHK(c, hipEventRecord(ev[2 * launch + k], c->stream));
    return 0;
  };
  for (int s = 0; s < nsteps; ++s) {
    const auto& q = *cur;
    const auto& o = *nxt;
    int rc = 0;
    if ((rc = rec(0)) || (rc = launch_stage_mixed<false, float>(c, q, nullptr, f1.m, f1.t, f1.u, dt / 4.0)) ||
        (rc = rec(1)) || (rc = launch_ghost_fill(c, f1.m, f1.t, f1.u)))
      return rc;
    ++launch;
    if ((rc = rec(0)) || (rc = launch_stage_mixed<true, float>(c, q, &f1, f2.m, f2.t, f2.u, dt / 3.0)) ||
        (rc = rec(1)) || (rc = launch_ghost_fill(c, f2.m, f2.t, f2.u)))
      return rc;
    ++launch;
    if ((rc = rec(0)) || (rc = launch_stage_mixed<true, float>(c, q, &f2, f1.m, f1.t, f1.u, dt / 2.0)) ||
        (rc = rec(1)) || (rc = launch_ghost_fill(c, f1.m, f1.t, f1.u)))
      return rc;
    ++launch;
    if ((rc = rec(0)) || (rc = launch_stage_mixed<true, double>(c, q, &f1, o.m, o.t, o.u, dt)) || (rc = rec(1)) ||
        (rc = launch_ghost_fill(c, o)))
      return rc;
    ++launch;
    std::swap(cur, nxt);
  }
  if (cur != &sol) {
    const size_t ns = (size_t)c->params.nlayers * c->nmembers * c->info.ncells;
    HIPCHK(c, hipMemcpyAsync(sol.m, cur->m, ns * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(sol.t, cur->t, ns * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(sol.u, cur->u, 3 * ns * 8, hipMemcpyDeviceToDevice, c->stream));
  }
  return 0;
}

// Set-up of RK4 steps on `slot`, outside any graph capture: the frozen halo
// slots of the stage buffers; in FP32-bulk mode also the FP32 buffers and the
// launch-shape tuning
int rk4_prepare(wtg_ctx* c, int slot, int nsteps) {
  if (int rc = check_slot(c, slot)) return rc;
  if (!c->topo_set) return fail(c, WTG_E_STATE, "wtg_set_topography must be called before time stepping");
  if (nsteps < 0) return fail(c, WTG_E_ARG, "nsteps < 0");
  const auto& sol = c->slots[slot];
  if (c->precision == WTG_PRECISION_FP32_BULK) {
    if (int rc = f32_ready(c)) return rc;
    if (nsteps == 0) return 0;
    if (!c->mixed_tuned)
      if (int rc = tune_mixed_shape(c, sol)) return rc;
    return launch_copy_frozen(c, sol, c->slots[c->nslots]);
  }
  if (int rc = launch_copy_frozen(c, sol, c->slots[c->nslots])) return rc;
  return launch_copy_frozen(c, sol, c->slots[c->nslots + 1]);
}

// nsteps RK4 steps (kernels only, after rk4_prepare); when ev is given,
// ev[2i] / ev[2i+1] are recorded around the i-th stage-kernel launch (no
// synchronisation between launches)
int rk4_steps(wtg_ctx* c, int slot, double dt, int nsteps, hipEvent_t* ev) {
  if (c->precision == WTG_PRECISION_FP32_BULK) return rk4_mixed_steps(c, slot, dt, nsteps, ev);
  const auto& sol = c->slots[slot];
  const auto& w1 = c->slots[c->nslots];
  const auto& w2 = c->slots[c->nslots + 1];
  const double coef[4] = {dt / 4.0, dt / 3.0, dt / 2.0, dt};
  int launch = 0;
  for (int s = 0; s < nsteps; ++s) {
    const wtg_ctx::Slot* ys[4] = {&sol, &w1, &w2, &w1};
    const wtg_ctx::Slot* os[4] = {&w1, &w2, &w1, &sol};
    for (int st = 0; st < 4; ++st, ++launch) {
      if (ev) HIPCHK(c, hipEventRecord(ev[2 * launch], c->stream));
      if (int rc = launch_stage(c, *ys[st], sol, *os[st], coef[st], 0)) return rc;
      if (ev) HIPCHK(c, hipEventRecord(ev[2 * launch + 1], c->stream));
      if (int rc = launch_ghost_fill(c, *os[st])) return rc;
    }
  }
  return 0;
}

int rk4_impl(wtg_ctx* c, int slot, double dt, int nsteps, hipEvent_t* ev) {
  if (int rc = rk4_prepare(c, slot, nsteps)) return rc;
  return rk4_steps(c, slot, dt, nsteps, ev);
}

// RK4 steps replayed from a captured HIP graph: one step (FP64) or two
// (FP32-bulk, whose state returns to the slot every second step) per graph;
// a remaining odd FP32-bulk step runs eagerly.  The graph holds kernel nodes
// only and is captured again when the slot, dt, precision or shape changes.
int rk4_graph_impl(wtg_ctx* c, int slot, double dt, int nsteps) {
  if (int rc = rk4_prepare(c, slot, nsteps)) return rc;
  const int per = c->precision == WTG_PRECISION_FP32_BULK ? 2 : 1;
  const int ng = nsteps / per, rest = nsteps % per;
  if (ng > 0) {
    const bool same = c->rk_graph && c->g_slot == slot && c->g_dt == dt && c->g_precision == c->precision &&
                      c->g_shape == c->mixed_shape;
    if (!same) {
      if (c->rk_graph) {
        hipGraphExecDestroy(c->rk_graph);
        c->rk_graph = nullptr;
      }
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = rk4_steps(c, slot, dt, per, nullptr);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
      }
      if (ec != hipSuccess) return fail(c, WTG_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
      const hipError_t ei = hipGraphInstantiate(&c->rk_graph, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        c->rk_graph = nullptr;
        return fail(c, WTG_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
      }
      c->g_slot = slot;
      c->g_dt = dt;
      c->g_precision = c->precision;
      c->g_shape = c->mixed_shape;
    }
    for (int i = 0; i < ng; ++i) HIPCHK(c, hipGraphLaunch(c->rk_graph, c->stream));
  }
  if (rest) return rk4_steps(c, slot, dt, rest, nullptr);
  return 0;
}


// RK4_split (time_integr.f90:360-382): four RK_split stages (444-462) and
// free_surface_update (464-478); the free surface solved by BiCGSTAB, the
// reference's SRJ on a single level
int baro_correction(wtg_ctx* c, const wtg_ctx::Slot& s) {
  const int n = c->nmaster;
  hipLaunchKernelGGL(baro_correction_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, c->d_ncl,
                     c->params.nlayers, c->info.ncells, c->params.ref_density, c->fs[FS_ETA], c->d_topo, s.m, s.t);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// the single-member views of an ensemble context for the free surface
// (built once; their scalar parameters follow the context's on every call)
int member_views(wtg_ctx* c, std::vector<wtg_ctx*>& out) {
  out.clear();
  if (c->nmembers == 1) {
    out.push_back(c);
    return 0;
  }
  if (c->views.empty()) {
    const int64_t NC = c->info.ncells, L = c->params.nlayers;
    std::vector<int64_t> src;
    src.swap(c->node_src);  // not needed by a view, and large
    for (int m = 0; m < c->nmembers; ++m) {
      wtg_ctx* v = new wtg_ctx(*c);
      v->allocs.clear();
      v->views.clear();
      v->own_stream = nullptr;
      v->rk_graph = nullptr;
      v->bicg_graph = nullptr;
      v->use_graph = false;
      v->fs_graph = false;
      v->nmembers = 1;
      v->d_topo = c->d_topo + m * NC;
      for (auto& sl : v->slots) {
        sl.m += m * L * NC;
        sl.t += m * L * NC;
        sl.u += 3 * m * L * NC;
      }
      for (int f = 0; f < FS_NFIELDS; ++f) v->fs[f] = c->fs[f] + m * NC;
      for (auto& f32 : v->fs32) f32 = nullptr;
      c->views.push_back(v);
    }
    src.swap(c->node_src);
  }
  const int64_t NC = c->info.ncells, L = c->params.nlayers;
  for (int m = 0; m < c->nmembers; ++m) {
    wtg_ctx* v = c->views[m];
    for (int k = 0; k < 2; ++k) {  // the FP32 buffers may be allocated after the views
      const auto& sf = c->fstage[k];
      v->fstage[k] = {sf.m ? sf.m + m * L * NC : nullptr, sf.t ? sf.t + m * L * NC : nullptr,
                      sf.u ? sf.u + 3 * m * L * NC : nullptr};
    }
    v->params = c->params;
    v->fs_dt = c->fs_dt;
    v->fs_theta1 = c->fs_theta1;
    v->fs_theta2 = c->fs_theta2;
    v->fs_params_set = c->fs_params_set;
    v->topo_set = c->topo_set;
    v->precision = c->precision;
    v->mixed_shape = c->mixed_shape;
    v->stream = c->stream;
    out.push_back(v);
  }
  return 0;
}

int bicgstab_solve(void* user, wtg_ctx* c, int* iters) {
  const double* tm = (const double*)user;  // {tol, maxit}
  double res = 0.0;
  return wtg_elliptic_solve(c, WTG_SOLVE_BICGSTAB, tm[0], (int)tm[1], iters, &res);
}

}  // namespace

int wtg_launch_stage_keep(wtg_ctx* c, const wtg_ctx::Slot& y, const wtg_ctx::Slot& q, const wtg_ctx::Slot& out,
                          double h, int trend_only, const WtgStageKeep* keep) {
  return launch_stage(c, y, q, out, h, trend_only, keep);
}

int wtg_ghost_fill_layers(wtg_ctx* c, int L, double* m, double* t, double* u) {
  const int nn = m ? c->nn : 0, ne = u ? c->ne : 0;
  const int64_t total = (int64_t)(nn + ne) * ((L + GF_KC - 1) / GF_KC);
  if (total == 0) return 0;
  const int threads = 256;
  const int64_t blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(ghost_fill_kernel<double>, dim3((unsigned)blocks), dim3(threads), 0, c->stream, L, c->info.ncells,
                     c->d_nd, nn, c->d_ed, c->d_es, ne, m, t, u);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int wtg_copy_frozen_slots(wtg_ctx* c, const wtg_ctx::Slot& from, const wtg_ctx::Slot& to) {
  return launch_copy_frozen(c, from, to);
}

int wtg_baro_correction_slot(wtg_ctx* c, const wtg_ctx::Slot& s) { return baro_correction(c, s); }

int wtg_split_axpy(wtg_ctx* c, double h, const double* eta_grad, const wtg_ctx::Slot& base, const wtg_ctx::Slot& tr,
                   const wtg_ctx::Slot& out) {
  const int L = c->params.nlayers;
  const unsigned nb = (unsigned)(((int64_t)c->nmaster * L + 255) / 256);
  hipLaunchKernelGGL(split_axpy_kernel<double>, dim3(nb), dim3(256), 0, c->stream, c->nmaster, c->d_ncl, L,
                     c->info.ncells, c->info.S, h, c->fs_theta1, c->params.grav, eta_grad, c->d_len, base.m, base.t,
                     base.u, tr.m, tr.t, tr.u, out.m, out.t, out.u);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int wtg_u_update(wtg_ctx* c, const double* eta_grad, const wtg_ctx::Slot& s) {
  const int L = c->params.nlayers;
  const unsigned nb = (unsigned)(((int64_t)c->nmaster * L + 255) / 256);
  hipLaunchKernelGGL(u_update_kernel, dim3(nb), dim3(256), 0, c->stream, c->nmaster, c->d_ncl, L, c->info.ncells,
                     c->info.S, c->fs_theta1 * c->fs_dt, c->params.grav, eta_grad, c->d_len, s.u);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int wtg_split_steps(wtg_ctx* c, int slot, int nsteps, wtg_fs_solver solve, void* user, int* iters, int nstages) {
  if (int rc = check_slot(c, slot)) return rc;
  if (nstages < 1 || nstages > 4) return fail(c, WTG_E_ARG, "split schemes have 1 to 4 stages, not %d", nstages);
  if (!c->topo_set) return fail(c, WTG_E_STATE, "wtg_set_topography must be called before time stepping");
  if (!c->fs_params_set) return fail(c, WTG_E_STATE, "wtg_fs_set_params must be called first");
  if (nsteps < 0) return fail(c, WTG_E_ARG, "negative step count");
  if (int rc = wtg_fs_ready(c)) return rc;
  const auto& sol = c->slots[slot];
  const auto& w1 = c->slots[c->nslots];
  const auto& w2 = c->slots[c->nslots + 1];
  const int L = c->params.nlayers;
  const int64_t NC = c->info.ncells, NCM = NC * c->nmembers;
  const double dt = c->fs_dt, th1 = c->fs_theta1, grav = c->params.grav;
  if (int rc = launch_copy_frozen(c, sol, w1)) return rc;
  if (int rc = launch_copy_frozen(c, sol, w2)) return rc;
  // FP32-bulk: the trend stencils run in FP32 into the first FP32 stage
  // buffer; the stage states, the AXPYs and the free surface stay FP64
  const bool f32 = c->precision == WTG_PRECISION_FP32_BULK;
  // FP64 with the layer-pair kernel: the stage's update is the split AXPY
  const bool fused = !f32 && c->stage_version == 3 && c->split_fused;
  if (f32) {
    if (int rc = f32_ready(c)) return rc;
    if (!c->mixed_tuned && nsteps > 0) {
      if (int rc = launch_ghost_fill(c, sol)) return rc;
      if (int rc = tune_mixed_shape(c, sol)) return rc;
    }
  }
  // one launch for the trend of every member and for each step of the
  // members' BiCGSTAB; the other 2-D and per-column steps on each member's view
  std::vector<wtg_ctx*> V;
  if (int rc = member_views(c, V)) return rc;
  auto view_err = [&](wtg_ctx* v, int rc) {
    if (rc && v != c) c->err = v->err;
    return rc;
  };
  const auto& ftr = c->fstage[0];
  // RK_split stages (time_integr.f90:360-451) with h/n, ..., h/2, h: sol ->
  // q1 -> ... -> sol, q1 alternating between the two work slots (a stage
  // kernel does not write the state it reads); the trend of the unfused path
  // goes to the stage's output slot, or the other work slot when that is sol
  double coef[4];
  int ins[4], trs[4], outs[4];
  for (int k = 0; k < nstages; ++k) {
    coef[k] = dt / (double)(nstages - k);
    outs[k] = k == nstages - 1 ? slot : (k % 2 == 0 ? c->nslots : c->nslots + 1);
    ins[k] = k == 0 ? slot : outs[k - 1];
    trs[k] = outs[k] != slot ? outs[k] : (ins[k] == c->nslots ? c->nslots + 1 : c->nslots);
  }
  const unsigned nb = (unsigned)(((int64_t)c->nmaster * L + 255) / 256);
  *iters = 0;
  for (int s = 0; s < nsteps; ++s) {
    if (int rc = launch_ghost_fill(c, sol)) return rc;
    // flux divergence of the layer-integrated velocity at the previous step
    for (wtg_ctx* v : V)
      if (int rc = wtg_fs_flux_div_async(v, slot)) return view_err(v, rc);
    HIPCHK(c, hipMemcpyAsync(c->fs[FS_DIVF_OLD], c->fs[FS_DIVF], NCM * sizeof(double), hipMemcpyDeviceToDevice,
                             c->stream));
    for (int st = 0; st < nstages; ++st) {
      const auto &in = c->slots[ins[st]], &tr = c->slots[trs[st]], &out = c->slots[outs[st]];
      for (wtg_ctx* v : V)
        if (int rc = baro_correction(v, v->slots[ins[st]])) return view_err(v, rc);
      if (int rc = launch_ghost_fill(c, in)) return rc;
      if (fused && ins[st] != outs[st]) {  // trend and u_star / scalar_star in one launch
        if (int rc = launch_stage(c, in, sol, out, coef[st], 0, nullptr, c->fs[FS_ETA])) return rc;
        if (int rc = launch_ghost_fill(c, out)) return rc;
        continue;
      }
      if (f32) {
        if (int rc = launch_stage_mixed<false, float>(c, in, nullptr, ftr.m, ftr.t, ftr.u, 1.0)) return rc;
      } else {
        if (int rc = launch_stage(c, in, in, tr, 0.0, 1)) return rc;
      }
      for (wtg_ctx* v : V) {
        const auto &vs = v->slots[slot], &vt = v->slots[trs[st]], &vo = v->slots[outs[st]];
        const auto& vf = v->fstage[0];
        if (f32)
          hipLaunchKernelGGL(split_axpy_kernel<float>, dim3(nb), dim3(256), 0, c->stream, c->nmaster, c->d_ncl, L, NC,
                             c->info.S, coef[st], th1, grav, v->fs[FS_ETA], c->d_len, vs.m, vs.t, vs.u, vf.m, vf.t,
                             vf.u, vo.m, vo.t, vo.u);
        else
          hipLaunchKernelGGL(split_axpy_kernel<double>, dim3(nb), dim3(256), 0, c->stream, c->nmaster, c->d_ncl, L,
                             NC, c->info.S, coef[st], th1, grav, v->fs[FS_ETA], c->d_len, vs.m, vs.t, vs.u, vt.m,
                             vt.t, vt.u, vo.m, vo.t, vo.u);
        HIPCHK(c, hipGetLastError());
      }
      if (int rc = launch_ghost_fill(c, out)) return rc;
    }
    // free_surface_update: eta_update, barotropic_correction, u_update
    for (wtg_ctx* v : V) {
      if (int rc = wtg_fs_flux_div_async(v, slot)) return view_err(v, rc);
      if (int rc = wtg_fs_rhs_async(v)) return view_err(v, rc);
    }
    HIPCHK(c, hipMemcpyAsync(c->fs[FS_ETA_OLD], c->fs[FS_ETA], NCM * sizeof(double), hipMemcpyDeviceToDevice,
                             c->stream));
    if (c->nmembers == 1 || solve != bicgstab_solve) {  // single member, or a member-batched hierarchy
      if (int rc = solve(user, c, iters)) return rc;
    } else {  // every member's BiCGSTAB in the same launches
      const double* tm = (const double*)user;  // {tol, maxit} of bicgstab_solve
      double err = 0.0;
      if (int rc = wtg_fs_bicgstab_all(c, tm[0], (int)tm[1], iters, &err)) return rc;
    }
    for (wtg_ctx* v : V) {
      if (int rc = baro_correction(v, v->slots[slot])) return view_err(v, rc);
      hipLaunchKernelGGL(u_update_kernel, dim3(nb), dim3(256), 0, c->stream, c->nmaster, c->d_ncl, L, NC, c->info.S,
                         th1 * dt, grav, v->fs[FS_ETA], c->d_len, v->slots[slot].u);
      HIPCHK(c, hipGetLastError());
    }
    if (int rc = launch_ghost_fill(c, sol)) return rc;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int wtg_grid_create(const wtg_grid_desc* desc, wtg_grid** out) {
  if (!desc || !out) return fail(nullptr, WTG_E_ARG, "null argument");
  *out = nullptr;
  try {
    auto* g = new wtg_grid;
    g->lv = wtg::build_level(desc->level, desc->min_level, desc->domain_level, desc->radius, desc->omega);
    if (g->lv.n % T != 0) {
      delete g;
      return fail(nullptr, WTG_E_ARG, "level %d: lozenge side %d is not a multiple of the tile %d", desc->level,
                  1 << desc->level, T);
    }
    *out = g;
  } catch (const std::exception& e) {
    return fail(nullptr, WTG_E_ARG, "wtg_grid_create: %s", e.what());
  }
  return 0;
}

int wtg_grid_info_get(const wtg_grid* grid, wtg_grid_info* info) {
  if (!grid || !info) return fail(nullptr, WTG_E_ARG, "null argument");
  info->n = grid->lv.n;
  info->S = grid->lv.S;
  info->halo = H;
  info->ncells = (int64_t)grid->lv.ncells();
  info->south_pole = (int64_t)grid->lv.south_pole;
  info->north_pole = (int64_t)grid->lv.north_pole;
  return 0;
}

int wtg_grid_get(const wtg_grid* grid, const char* name, void* out, size_t bytes) {
  if (!grid || !name || !out) return fail(nullptr, WTG_E_ARG, "null argument");
  const auto& g = grid->lv;
  const void* src = nullptr;
  size_t n = 0;
  std::string s(name);
  if (s == "x") { src = g.X.data(); n = g.X.size() * 8; }
  else if (s == "len") { src = g.len.data(); n = g.len.size() * 8; }
  else if (s == "pedlen") { src = g.pedlen.data(); n = g.pedlen.size() * 8; }
  else if (s == "area") { src = g.area.data(); n = g.area.size() * 8; }
  else if (s == "hex_inv") { src = g.hex_inv.data(); n = g.hex_inv.size() * 8; }
  else if (s == "triarea") { src = g.triarea.data(); n = g.triarea.size() * 8; }
  else if (s == "coriolis") { src = g.coriolis.data(); n = g.coriolis.size() * 8; }
  else if (s == "node_src") { src = g.node_src.data(); n = g.node_src.size() * 8; }
  else if (s == "edge_src") { src = g.edge_src.data(); n = g.edge_src.size() * 8; }
  else if (s == "edge_sign") { src = g.edge_sign.data(); n = g.edge_sign.size(); }
  else if (s == "red_weight") { src = g.red_weight.data(); n = g.red_weight.size() * 8; }
  else if (s == "wide_src") { src = g.wide_src.data(); n = g.wide_src.size() * 8; }
  else return fail(nullptr, WTG_E_ARG, "unknown grid array '%s'", name);
  if (bytes != n) return fail(nullptr, WTG_E_ARG, "grid array '%s' has %zu bytes, %zu given", name, n, bytes);
  std::memcpy(out, src, n);
  return 0;
}

int wtg_grid_overlap_areas(const wtg_grid* fine, const wtg_grid* coarse, double* out, size_t bytes) {
  if (!fine || !coarse || !out) return fail(nullptr, WTG_E_ARG, "null argument");
  if (fine->lv.level != coarse->lv.level + 1) return fail(nullptr, WTG_E_ARG, "grids are not consecutive levels");
  const size_t n = 4 * fine->lv.ncells() * sizeof(double);
  if (bytes != n) return fail(nullptr, WTG_E_ARG, "overlap areas have %zu bytes, %zu given", n, bytes);
  try {
    std::vector<double> a = wtg::overlap_areas(fine->lv, coarse->lv);
    std::memcpy(out, a.data(), n);
  } catch (const std::exception& e) {
    return fail(nullptr, WTG_E_ARG, "%s", e.what());
  }
  return 0;
}

int wtg_grid_transfer_weights(const wtg_grid* fine, const wtg_grid* coarse, const char* name, double* out,
                              size_t bytes) {
  if (!fine || !coarse || !name || !out) return fail(nullptr, WTG_E_ARG, "null argument");
  if (fine->lv.level != coarse->lv.level + 1) return fail(nullptr, WTG_E_ARG, "grids are not consecutive levels");
  try {
    wtg::TransferWeights w = wtg::transfer_weights(fine->lv, coarse->lv);
    const std::vector<double>* v = nullptr;
    if (!std::strcmp(name, "split")) v = &w.split;
    else if (!std::strcmp(name, "iu")) v = &w.iu;
    else if (!std::strcmp(name, "rf")) v = &w.rf;
    else return fail(nullptr, WTG_E_ARG, "unknown transfer weights '%s'", name);
    const size_t n = v->size() * sizeof(double);
    if (bytes != n) return fail(nullptr, WTG_E_ARG, "transfer weights '%s' have %zu bytes, %zu given", name, n, bytes);
    std::memcpy(out, v->data(), n);
  } catch (const std::exception& e) {
    return fail(nullptr, WTG_E_ARG, "%s", e.what());
  }
  return 0;
}

// ghost entries of a uniform level by whether their source lies in the same
// shard as their destination, shards being whole lozenges: lozenge lz goes to
// shard lz * nshards / 10 (the reference's domains of one lozenge together)
int wtg_grid_shard_halo(const wtg_grid* grid, int nshards, int64_t counts[4]) {
  if (!grid || !counts) return fail(nullptr, WTG_E_ARG, "null argument");
  if (nshards < 1 || nshards > 10) return fail(nullptr, WTG_E_ARG, "nshards %d out of range [1,10]", nshards);
  const auto& g = grid->lv;
  const int64_t NC = (int64_t)g.ncells(), SS = (int64_t)g.S * g.S;
  const int n = g.n;
  auto shard = [&](int64_t cell) { return (int)(cell / SS) * nshards / 10; };
  for (int q = 0; q < 4; ++q) counts[q] = 0;
  for (int64_t c = 0; c < NC; ++c) {
    const int64_t r = c % SS;
    const int j = (int)(r / g.S) - H, i = (int)(r % g.S) - H;
    if (i >= 0 && i < n && j >= 0 && j < n) continue;
    const int64_t s = g.node_src[c];
    if (s >= 0 && s != c) ++counts[shard(s) == shard(c) ? 0 : 1];
    for (int e = 0; e < 3; ++e) {
      const int64_t slot = 3 * c + e, se = g.edge_src[slot];
      if (se >= 0 && se != slot) ++counts[shard(se / 3) == shard(c) ? 2 : 3];
    }
  }
  return 0;
}

void wtg_grid_destroy(wtg_grid* grid) { delete grid; }

int wtg_create(const wtg_grid* grid, const wtg_params* params, int nslots, wtg_ctx** out) {
  return wtg_create_ensemble(grid, params, 1, nslots, out);
}

int wtg_create_ensemble(const wtg_grid* grid, const wtg_params* params, int nmembers, int nslots, wtg_ctx** out) {
  if (!grid || !params || !out) return fail(nullptr, WTG_E_ARG, "null argument");
  *out = nullptr;
  if (nslots < 1 || nslots > 64) return fail(nullptr, WTG_E_ARG, "nslots %d out of range [1,64]", nslots);
  if (nmembers < 1 || nmembers > 1024) return fail(nullptr, WTG_E_ARG, "nmembers %d out of range [1,1024]", nmembers);
  if (params->nlayers < 1) return fail(nullptr, WTG_E_ARG, "nlayers %d < 1", params->nlayers);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, WTG_E_DEVICE, "no HIP device");
  std::unique_ptr<wtg_ctx> c(new wtg_ctx);
  if (hipGetDevice(&c->device) != hipSuccess) return fail(nullptr, WTG_E_DEVICE, "hipGetDevice failed");
  if (hipStreamCreate(&c->stream) != hipSuccess) return fail(nullptr, WTG_E_DEVICE, "hipStreamCreate failed");
  c->own_stream = c->stream;
  if (wtg_noalloc_scratch_init(c->device) != 0) {
    hipStreamDestroy(c->stream);
    return fail(nullptr, WTG_E_DEVICE, "hipGetSymbolAddress of the dot-product scratch failed on device %d", c->device);
  }
  wtg_grid_info_get(grid, &c->info);
  c->params = *params;
  c->nslots = nslots;
  c->nmembers = nmembers;
  if (const char* v = std::getenv("WTG_STAGE_KERNEL")) {
    int sv = std::atoi(v);
    c->stage_version = (sv == 1 || sv == 2) ? sv : 3;
  }
  // WTG_SPLIT_FUSED=0: the split-explicit stage as a trend launch and a
  // separate AXPY (the two give the same bits)
  if (const char* v = std::getenv("WTG_SPLIT_FUSED")) c->split_fused = std::atoi(v) != 0;
  const auto& g = grid->lv;
  const int64_t NC = c->info.ncells;
  // flux dual lengths: a halo edge that is ghost filled carries its owner's flux
  std::vector<double> pedf(3 * NC);
  for (int64_t cc = 0; cc < NC; ++cc)
    for (int e = 0; e < 3; ++e) {
      int64_t slot = 3 * cc + e, s = g.edge_src[slot];
      pedf[e * NC + cc] = (s >= 0 && s != slot) ? g.pedlen[(s % 3) * NC + s / 3] : g.pedlen[e * NC + cc];
    }
  // ghost lists and frozen lists
  std::vector<int64_t> nd, ed, fn, fe;
  std::vector<int8_t> es;
  const int n = c->info.n, S = c->info.S;
  const int n_ = n, S_ = S;
  for (int64_t cc = 0; cc < NC; ++cc) {
    int64_t rr = cc % ((int64_t)S * S);
    int j = (int)(rr / S) - H, i = (int)(rr % S) - H;
    bool owned = (i >= 0 && i < n && j >= 0 && j < n);
    if (owned) continue;
    int64_t s = g.node_src[cc];
    if (s >= 0 && s != cc) { nd.push_back(cc); nd.push_back(s); }
    if (s < 0) fn.push_back(cc);
    for (int e = 0; e < 3; ++e) {
      int64_t slot = 3 * cc + e, se = g.edge_src[slot];
      if (se >= 0 && se != slot) {
        ed.push_back(e * NC + cc);
        ed.push_back((se % 3) * NC + se / 3);
        es.push_back(g.edge_sign[slot]);
      }
      if (se < 0) fe.push_back(e * NC + cc);
    }
  }
  // free-surface problem: flux lengths, node compute list, reduction weights
  std::vector<double> lenf(3 * NC);
  for (int64_t cc = 0; cc < NC; ++cc)
    for (int e = 0; e < 3; ++e) {
      int64_t slot = 3 * cc + e, s = g.edge_src[slot];
      lenf[e * NC + cc] = (s >= 0 && s != slot) ? g.len[(s % 3) * NC + s / 3] : g.len[e * NC + cc];
    }
  std::vector<int64_t> ncl, rcell;
  std::vector<double> rw;
  for (int64_t cc = 0; cc < NC; ++cc) {
    int64_t rr = cc % ((int64_t)S_ * S_);
    int j = (int)(rr / S_) - H, i = (int)(rr % S_) - H;
    if ((i >= 0 && i < n_ && j >= 0 && j < n_) || cc == (int64_t)g.south_pole || cc == (int64_t)g.north_pole) {
      ncl.push_back(cc);
      ncl.push_back(cc);
    }
    if (g.red_weight[cc] > 0) {
      rcell.push_back(cc);
      rw.push_back(g.red_weight[cc]);
    }
  }
  c->nmaster = (int)(ncl.size() / 2);
  // ghost entries in destination order: the entries are independent copies,
  // and neighbouring threads of the fill then touch neighbouring halo cells
  {
    std::vector<size_t> o(nd.size() / 2);
    for (size_t q = 0; q < o.size(); ++q) o[q] = q;
    std::sort(o.begin(), o.end(), [&](size_t a, size_t b) { return nd[2 * a] < nd[2 * b]; });
    std::vector<int64_t> nd2(nd.size());
    for (size_t q = 0; q < o.size(); ++q) nd2[2 * q] = nd[2 * o[q]], nd2[2 * q + 1] = nd[2 * o[q] + 1];
    nd.swap(nd2);
    std::vector<size_t> oe(ed.size() / 2);
    for (size_t q = 0; q < oe.size(); ++q) oe[q] = q;
    std::sort(oe.begin(), oe.end(), [&](size_t a, size_t b) { return ed[2 * a] < ed[2 * b]; });
    std::vector<int64_t> ed2(ed.size());
    std::vector<int8_t> es2(es.size());
    for (size_t q = 0; q < oe.size(); ++q) {
      ed2[2 * q] = ed[2 * oe[q]], ed2[2 * q + 1] = ed[2 * oe[q] + 1];
      es2[q] = es[oe[q]];
    }
    ed.swap(ed2);
    es.swap(es2);
  }
  for (size_t q = 0; q < nd.size(); q += 2) {
    ncl.push_back(nd[q]);
    ncl.push_back(nd[q + 1]);
  }
  c->nncl = (int)(ncl.size() / 2);
  c->nred = (int)rcell.size();
  c->nn = (int)(nd.size() / 2);
  c->ne = (int)(ed.size() / 2);
  c->nfn = (int)fn.size();
  c->nfe = (int)fe.size();
  c->node_src = g.node_src;
  wtg_ctx* cp = c.get();
  int rc = 0;
  auto up = [&](auto** dptr, const auto* host, size_t count) -> int {
    using Tp = std::remove_const_t<std::remove_reference_t<decltype(*host)>>;
    if (int r2 = dalloc<Tp>(cp, (Tp**)dptr, count)) return r2;
    if (count) HIPCHK(cp, hipMemcpy(*dptr, host, count * sizeof(Tp), hipMemcpyHostToDevice));
    return 0;
  };
  if ((rc = up(&c->d_len, g.len.data(), g.len.size())) || (rc = up(&c->d_ped, g.pedlen.data(), g.pedlen.size())) ||
      (rc = up(&c->d_pedf, pedf.data(), pedf.size())) || (rc = up(&c->d_area, g.area.data(), g.area.size())) ||
      (rc = up(&c->d_hinv, g.hex_inv.data(), g.hex_inv.size())) ||
      (rc = up(&c->d_tri, g.triarea.data(), g.triarea.size())) ||
      (rc = up(&c->d_cor, g.coriolis.data(), g.coriolis.size())) || (rc = up(&c->d_nd, nd.data(), nd.size())) ||
      (rc = up(&c->d_ed, ed.data(), ed.size())) || (rc = up(&c->d_es, es.data(), es.size())) ||
      (rc = up(&c->d_fn, fn.data(), fn.size())) || (rc = up(&c->d_fe, fe.data(), fe.size())) ||
      (rc = up(&c->d_lenf, lenf.data(), lenf.size())) || (rc = up(&c->d_ncl, ncl.data(), ncl.size())) ||
      (rc = up(&c->d_rcell, rcell.data(), rcell.size())) || (rc = up(&c->d_rw, rw.data(), rw.size()))) {
    g_create_error = c->err;
    wtg_destroy(c.release());
    return rc;
  }
  if ((rc = dalloc<double>(cp, &c->d_topo, (size_t)nmembers * NC))) {
    g_create_error = c->err;
    wtg_destroy(c.release());
    return rc;
  }
  const int L = params->nlayers * c->nmembers;  // layers of all ensemble members
  c->slots.resize(nslots + 2);
  for (auto& s : c->slots) {
    if ((rc = dalloc<double>(cp, &s.m, (size_t)L * NC)) || (rc = dalloc<double>(cp, &s.t, (size_t)L * NC)) ||
        (rc = dalloc<double>(cp, &s.u, (size_t)3 * L * NC))) {
      g_create_error = c->err;
      wtg_destroy(c.release());
      return rc;
    }
    hipMemset(s.m, 0, sizeof(double) * L * NC);
    hipMemset(s.t, 0, sizeof(double) * L * NC);
    hipMemset(s.u, 0, sizeof(double) * 3 * L * NC);
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    g_create_error = "device synchronisation after allocation failed";
    wtg_destroy(c.release());
    return WTG_E_DEVICE;
  }
  *out = c.release();
  return 0;
}

void wtg_destroy(wtg_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->rk_graph) hipGraphExecDestroy(c->rk_graph);
  if (c->bicg_graph) hipGraphExecDestroy(c->bicg_graph);
  for (wtg_ctx* v : c->views) delete v;  // views own no device memory
  for (void* p : c->allocs) hipFree(p);
  if (c->h_scal) hipHostFree(c->h_scal);
  // the context's own stream (a hierarchy may have pointed c->stream at its
  // finest level's, which that level destroys)
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
}

int wtg_set_topography(wtg_ctx* c, const double* topo) {
  if (!c || !topo) return WTG_E_ARG;
  const int64_t NC = c->info.ncells, n = (int64_t)c->nmembers * NC;
  std::vector<double> t(topo, topo + n);
  for (int64_t m = 0; m < n; m += NC)
    for (int64_t cc = 0; cc < NC; ++cc) {
      int64_t s = c->node_src[cc];
      if (s >= 0) t[m + cc] = topo[m + s];
    }
  HIPCHK(c, hipMemcpyAsync(c->d_topo, t.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->topo_set = true;
  return 0;
}

int wtg_upload_state(wtg_ctx* c, int slot, const double* mass, const double* temp, const double* velo) {
  if (int rc = check_slot(c, slot)) return rc;
  if (!mass || !temp || !velo) return fail(c, WTG_E_ARG, "null state array");
  const size_t ns = (size_t)c->params.nlayers * c->nmembers * c->info.ncells;
  const auto& s = c->slots[slot];
  HIPCHK(c, hipMemcpyAsync(s.m, mass, ns * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.t, temp, ns * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.u, velo, 3 * ns * 8, hipMemcpyHostToDevice, c->stream));
  if (int rc = launch_ghost_fill(c, s)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_download_state(wtg_ctx* c, int slot, double* mass, double* temp, double* velo) {
  if (int rc = check_slot(c, slot)) return rc;
  if (!mass || !temp || !velo) return fail(c, WTG_E_ARG, "null state array");
  const size_t ns = (size_t)c->params.nlayers * c->nmembers * c->info.ncells;
  const auto& s = c->slots[slot];
  HIPCHK(c, hipMemcpyAsync(mass, s.m, ns * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(temp, s.t, ns * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(velo, s.u, 3 * ns * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// the ghost fill of `slot` as two launches, the entries whose source lies in
// the destination's shard and those whose source does not (whole-lozenge
// shards as wtg_grid_shard_halo); each timed over `reps` launches
int wtg_ghost_fill_split_timed(wtg_ctx* c, int slot, int nshards, int reps, double* ms_local, double* ms_remote) {
  if (int rc = check_slot(c, slot)) return rc;
  if (!ms_local || !ms_remote) return fail(c, WTG_E_ARG, "null argument");
  if (nshards < 1 || nshards > 10) return fail(c, WTG_E_ARG, "nshards %d out of range [1,10]", nshards);
  if (reps < 1 || reps > 10000) return fail(c, WTG_E_ARG, "reps %d out of range [1,10000]", reps);
  const int64_t NC = c->info.ncells, SS = (int64_t)c->info.S * c->info.S;
  auto shard = [&](int64_t cell) { return (int)(cell / SS) * nshards / 10; };
  std::vector<int64_t> nd(2 * (size_t)c->nn), ed(2 * (size_t)c->ne);
  std::vector<int8_t> es(c->ne);
  HIPCHK(c, hipMemcpy(nd.data(), c->d_nd, nd.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ed.data(), c->d_ed, ed.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(es.data(), c->d_es, es.size(), hipMemcpyDeviceToHost));
  std::vector<int64_t> ndp[2], edp[2];
  std::vector<int8_t> esp[2];
  for (int q = 0; q < c->nn; ++q) {
    const int r = shard(nd[2 * q]) == shard(nd[2 * q + 1]) ? 0 : 1;
    ndp[r].push_back(nd[2 * q]);
    ndp[r].push_back(nd[2 * q + 1]);
  }
  for (int q = 0; q < c->ne; ++q) {  // plane indices e * NC + cell
    const int r = shard(ed[2 * q] % NC) == shard(ed[2 * q + 1] % NC) ? 0 : 1;
    edp[r].push_back(ed[2 * q]);
    edp[r].push_back(ed[2 * q + 1]);
    esp[r].push_back(es[q]);
  }
  const auto& s = c->slots[slot];
  const int L = c->params.nlayers * c->nmembers;
  double ms[2] = {0.0, 0.0};
  int rc = 0;
  for (int r = 0; r < 2 && !rc; ++r) {
    int64_t *dn = nullptr, *de = nullptr;
    int8_t* dsg = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int nn = (int)(ndp[r].size() / 2), ne = (int)(edp[r].size() / 2);
    if (hipMalloc(&dn, std::max<size_t>(ndp[r].size(), 1) * 8) != hipSuccess ||
        hipMalloc(&de, std::max<size_t>(edp[r].size(), 1) * 8) != hipSuccess ||
        hipMalloc(&dsg, std::max<size_t>(esp[r].size(), 1)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
      rc = fail(c, WTG_E_DEVICE, "allocation of the split lists failed");
    if (!rc && (hipMemcpy(dn, ndp[r].data(), ndp[r].size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(de, edp[r].data(), edp[r].size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dsg, esp[r].data(), esp[r].size(), hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(c, WTG_E_DEVICE, "upload of the split lists failed");
    const int64_t total = (int64_t)(nn + ne) * ((L + GF_KC - 1) / GF_KC);
    if (!rc && total > 0) {
      const unsigned blocks = (unsigned)((total + 255) / 256);
      hipEventRecord(e0, c->stream);
      for (int k = 0; k < reps; ++k)
        hipLaunchKernelGGL(ghost_fill_kernel<double>, dim3(blocks), dim3(256), 0, c->stream, L, NC, dn, nn, de, dsg,
                           ne, s.m, s.t, s.u);
      hipEventRecord(e1, c->stream);
      if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        rc = fail(c, WTG_E_DEVICE, "split ghost fill failed");
      float t = 0.f;
      if (!rc) hipEventElapsedTime(&t, e0, e1);
      ms[r] = t / reps;
    }
    hipStreamSynchronize(c->stream);
    if (dn) hipFree(dn);
    if (de) hipFree(de);
    if (dsg) hipFree(dsg);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
  }
  *ms_local = ms[0];
  *ms_remote = ms[1];
  return rc;
}

int wtg_ghost_fill(wtg_ctx* c, int slot) {
  if (int rc = check_slot(c, slot)) return rc;
  if (int rc = launch_ghost_fill(c, c->slots[slot])) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_trend(wtg_ctx* c, int q_slot, int dq_slot) {
  if (int rc = check_slot(c, q_slot)) return rc;
  if (int rc = check_slot(c, dq_slot)) return rc;
  if (q_slot == dq_slot) return fail(c, WTG_E_ARG, "q_slot and dq_slot must differ");
  if (!c->topo_set) return fail(c, WTG_E_STATE, "wtg_set_topography must be called before wtg_trend");
  const size_t ns = (size_t)c->params.nlayers * c->nmembers * c->info.ncells;
  const auto& dq = c->slots[dq_slot];
  HIPCHK(c, hipMemsetAsync(dq.m, 0, ns * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(dq.t, 0, ns * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(dq.u, 0, 3 * ns * 8, c->stream));
  const auto& q = c->slots[q_slot];
  if (int rc = launch_stage(c, q, q, dq, 0.0, 1)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_rk_split(wtg_ctx* c, int slot, int nstages, int nsteps, double rel_tol, int max_iter, int* iters) {
  if (!c || !iters) return WTG_E_ARG;
  *iters = 0;
  if (max_iter < 0) return fail(c, WTG_E_ARG, "max_iter < 0");
  double tm[2] = {rel_tol, (double)max_iter};
  if (int rc = wtg_split_steps(c, slot, nsteps, bicgstab_solve, tm, iters, nstages)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_rk4_split(wtg_ctx* c, int slot, int nsteps, double rel_tol, int max_iter, int* iters) {
  return wtg_rk_split(c, slot, 4, nsteps, rel_tol, max_iter, iters);
}

int wtg_restrict_velo(wtg_ctx* fine, int fine_slot, wtg_ctx* coarse, int coarse_slot) {
  if (!fine || !coarse) return WTG_E_ARG;
  if (int rc = check_slot(fine, fine_slot)) return rc;
  if (int rc = check_slot(coarse, coarse_slot)) return fail(fine, rc, "coarse slot: %s", coarse->err.c_str());
  if (coarse->info.n * 2 != fine->info.n) return fail(fine, WTG_E_ARG, "the coarse level is not the next coarser one");
  const int L = fine->params.nlayers * fine->nmembers;
  if (coarse->params.nlayers * coarse->nmembers != L) return fail(fine, WTG_E_ARG, "the levels hold different layers");
  if (coarse->device != fine->device) return fail(fine, WTG_E_ARG, "the levels are on different devices");
  const int nc = coarse->info.n;
  const int64_t total = (int64_t)10 * nc * nc * L;
  // on the coarse level's stream, after the fine level's work
  HIPCHK(fine, hipStreamSynchronize(fine->stream));
  hipLaunchKernelGGL(restrict_velo_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, coarse->stream, nc,
                     coarse->info.S, fine->info.S, coarse->info.ncells, fine->info.ncells, L,
                     fine->slots[fine_slot].u, coarse->slots[coarse_slot].u);
  HIPCHK(fine, hipGetLastError());
  if (int rc = launch_ghost_fill(coarse, coarse->slots[coarse_slot])) return fail(fine, rc, "%s", coarse->err.c_str());
  HIPCHK(fine, hipStreamSynchronize(coarse->stream));
  return 0;
}

int wtg_set_graph(wtg_ctx* c, int enable) {
  if (!c) return WTG_E_ARG;
  c->use_graph = enable != 0;
  return 0;
}

int wtg_rk4(wtg_ctx* c, int slot, double dt, int nsteps) {
  if (!c) return WTG_E_ARG;
  if (int rc = c->use_graph ? rk4_graph_impl(c, slot, dt, nsteps) : rk4_impl(c, slot, dt, nsteps, nullptr)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_rk4_timed(wtg_ctx* c, int slot, double dt, int nsteps, double* total_ms, double* stage_kernel_ms) {
  if (!c || !total_ms || !stage_kernel_ms) return WTG_E_ARG;
  if (nsteps < 1 || nsteps > 100000) return fail(c, WTG_E_ARG, "nsteps %d out of range [1,100000]", nsteps);
  const int nl = 4 * nsteps;
  std::vector<hipEvent_t> ev(2 * nl + 2, nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { rc = fail(c, WTG_E_DEVICE, "hipEventCreate failed"); break; }
  if (!rc) {
    // whole job between ev[2nl] and ev[2nl+1]; every stage launch between its own pair
    if (hipStreamSynchronize(c->stream) != hipSuccess || hipEventRecord(ev[2 * nl], c->stream) != hipSuccess)
      rc = fail(c, WTG_E_DEVICE, "event record failed");
    if (!rc) rc = rk4_impl(c, slot, dt, nsteps, ev.data());
    if (!rc && (hipEventRecord(ev[2 * nl + 1], c->stream) != hipSuccess ||
                hipEventSynchronize(ev[2 * nl + 1]) != hipSuccess))
      rc = fail(c, WTG_E_DEVICE, "event synchronise failed");
    if (!rc) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev[2 * nl], ev[2 * nl + 1]);
      *total_ms = ms;
      double sum = 0.0;
      for (int i = 0; i < nl; ++i) {
        float k = 0.f;
        hipEventElapsedTime(&k, ev[2 * i], ev[2 * i + 1]);
        sum += k;
      }
      *stage_kernel_ms = sum / nl;
    }
  }
  for (auto& e : ev)
    if (e) hipEventDestroy(e);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_set_precision(wtg_ctx* c, int precision) {
  if (!c) return WTG_E_ARG;
  if (precision != WTG_PRECISION_FP64 && precision != WTG_PRECISION_FP32_BULK)
    return fail(c, WTG_E_ARG, "unknown precision mode %d", precision);
  if (precision == WTG_PRECISION_FP32_BULK && c->dif_on)
    return fail(c, WTG_E_STATE, "diffusion is FP64 only; switch it off before the FP32-bulk precision");
  if (precision == WTG_PRECISION_FP32_BULK)
    if (int rc = f32_ready(c)) return rc;
  c->precision = precision;
  return 0;
}

int wtg_set_diffusion(wtg_ctx* c, const wtg_diffusion* d) {
  if (!c) return WTG_E_ARG;
  const wtg_diffusion off{};
  const wtg_diffusion& v = d ? *d : off;
  for (int o : {v.laplace_sclr, v.laplace_divu, v.laplace_rotu})
    if (o < 0 || o > 2) return fail(c, WTG_E_ARG, "Laplacian orders are 0, 1 or 2, not %d", o);
  if (v.laplace_divu && !v.laplace_rotu)
    return fail(c, WTG_E_ARG, "laplace_divu needs laplace_rotu: the velocity term's sign is (-1)^(laplace_rotu-1)");
  const bool on = v.laplace_sclr || v.laplace_rotu;
  if (on && c->precision != WTG_PRECISION_FP64)
    return fail(c, WTG_E_STATE, "diffusion is FP64 only (the context is in the FP32-bulk precision)");
  if (on && c->stage_version != 3)
    return fail(c, WTG_E_STATE, "diffusion needs the layer-pair stage kernel (WTG_STAGE_KERNEL unset)");
  c->dif = v;
  c->dif_on = on;
  if (on)
    if (int rc = wtg_diffusion_alloc(c)) {
      c->dif_on = false;
      return rc;
    }
  // a captured RK4 graph holds the previous settings' launches
  if (c->rk_graph) {
    hipGraphExecDestroy(c->rk_graph);
    c->rk_graph = nullptr;
  }
  return 0;
}

int wtg_set_stage_shape(wtg_ctx* c, int shape) {
  if (!c) return WTG_E_ARG;
  if (shape == WTG_SHAPE_AUTO) {
    c->mixed_tuned = false;
    c->mixed_shape = WTG_SHAPE_LAYER_PAIRS;
    return 0;
  }
  if (shape != WTG_SHAPE_LAYER_PAIRS && shape != WTG_SHAPE_EDGE_TASKS && shape != WTG_SHAPE_ONE_LAYER)
    return fail(c, WTG_E_ARG, "unknown stage shape %d", shape);
  c->mixed_shape = shape;
  c->mixed_tuned = true;
  c->tuned_ms[0] = c->tuned_ms[1] = c->tuned_ms[2] = 0.f;
  return 0;
}

int wtg_get_stage_shape(wtg_ctx* c, int* shape, float* ms) {
  if (!c || !shape) return WTG_E_ARG;
  *shape = c->mixed_tuned ? c->mixed_shape : WTG_SHAPE_AUTO;
  if (ms)
    for (int i = 0; i < 3; ++i) ms[i] = c->tuned_ms[i];
  return 0;
}

int wtg_state_device_ptrs(wtg_ctx* c, int slot, double** mass, double** temp, double** velo) {
  if (int rc = check_slot(c, slot)) return rc;
  if (mass) *mass = c->slots[slot].m;
  if (temp) *temp = c->slots[slot].t;
  if (velo) *velo = c->slots[slot].u;
  return 0;
}

int wtg_synchronize(wtg_ctx* c) {
  if (!c) return WTG_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_last_error(const wtg_ctx* c, char* buf, size_t n) {
  if (!buf || n == 0) return WTG_E_ARG;
  const std::string& s = c ? c->err : g_create_error;
  std::snprintf(buf, n, "%s", s.c_str());
  return 0;
}

}  // extern "C"
