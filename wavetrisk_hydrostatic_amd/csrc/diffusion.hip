// Laplacian diffusion and hyperdiffusion of the TRiSK trend (wtg_set_diffusion).
//
// The reference switches these on with Laplace_sclr / Laplace_divu /
// Laplace_rotu (shared.f90:226, 238) and applies them through the test
// cases' physics hooks inside step1 and du_source (ops.f90:360, 511, 778):
//   mass / temperature flux += (-1)^Ls visc grad(q or Laplacian q) pedlen
//   velocity source          += (-1)^(Lr-1) (visc_divu grad D - visc_rotu curl W)
// with D = div u (Ld = 1) or its Laplacian (Ld = 2) and W the vorticity
// (Lr = 1) or rot of rot of it (Lr = 2).  Before each stage kernel the
// pre-pass below evaluates, for every layer of the stage's input state:
//   cal_divu_ml            div(u pedlen) with cal_div over i, j in [0, n] of
//                          each lozenge and no ghost fill after it
//                          (multi_level.f90:579-617)
//   cal_Laplacian_divu,    div(grad q pedlen): gradient fluxes on the cells
//   cal_Laplacian_scalars  [-1, n]^2, halo edges then take their owners'
//                          fluxes (update_bdry of horiz_flux), divergence on
//                          the masters, then the halo (multi_level.f90:190-309)
//   vorticity              circulation / triangle area on the triangles of
//                          cells [-1, n]^2 with the four pentagon corners'
//                          merged triangles (step1 itype 0 / 8, post_step1,
//                          post_vort; ops.f90:256-275, 334-338, 1159-1278)
//   cal_Laplacian_vector_rot  rot u = -(vorticity difference) / pedlen on the
//                          owned edges, ghost filled, and its vorticity
//                          (multi_level.f90:240-270)
// and writes the hooks' values: the scalar diffusive fluxes on every edge of
// the cells [-1, n]^2 and the velocity source on the owned edges, each
// [layer][3][ncells].  The stage kernel adds them where the hooks enter the
// reference's expressions (flux = u pedlen interp(q) + flux_phys; source =
// -Qperp + phys len), so the FP64 step matches the oracle bit for bit.
//
// Each kernel is one thread per (cell, layer) over a square of lozenge cells
// or a list, cells fastest: HBM-bound stencils of a few neighbours.
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace {

constexpr int H = wtg::HALO;
constexpr int DNT = 256;

struct DGeo {
  const double *len, *ped, *lenf, *pedf, *hinv, *tri;
  int64_t NC;
  int n, S;
};

struct DPar {
  int ls, ld, lr;
  double vm, vt, vd, vr;
};

// thread -> (layer k, lozenge lz, cell (i, j)) of the square [lo, lo + side)^2
// of every lozenge; false past the end
__device__ __forceinline__ bool square(const DGeo& g, int lo, int side, int LM, int& k, int& i, int& j,
                                       int64_t& c) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t sq = (int64_t)side * side, per = 10 * sq;
  if (idx >= per * LM) return false;
  k = (int)(idx / per);
  int64_t r = idx % per;
  const int lz = (int)(r / sq);
  r %= sq;
  j = lo + (int)(r / side);
  i = lo + (int)(r % side);
  c = (int64_t)lz * g.S * g.S + (int64_t)(j + H) * g.S + (i + H);
  return true;
}

unsigned nblocks(int64_t n) { return (unsigned)((n + DNT - 1) / DNT); }

// cal_divu_ml: div of u pedlen (flux lengths: the owner's on halo edges) on
// i, j in [0, n], each lozenge's own value on its halo row and column
__global__ void dif_divu_kernel(DGeo g, int LM, const double* __restrict__ u, double* __restrict__ divu) {
  int k, i, j;
  int64_t c;
  if (!square(g, 0, g.n + 1, LM, k, i, j, c)) return;
  const int64_t NC = g.NC, S = g.S;
  const double* uk = u + (int64_t)k * 3 * NC;
#define FX(e, x) (uk[(int64_t)(e) * NC + (x)] * g.pedf[(int64_t)(e) * NC + (x)])
  divu[(int64_t)k * NC + c] =
      (FX(0, c) - FX(0, c - 1) + FX(1, c - S - 1) - FX(1, c) + FX(2, c) - FX(2, c - S)) * g.hinv[c];
#undef FX
}

// halo copies of a node field, f[dst] = f[src] per layer; outer: only the
// halo cells outside [0, n]^2 (the others keep the lozenge's own value)
__global__ void dif_node_fill_kernel(int LM, int64_t NC, int n, int S, const int64_t* __restrict__ nd, int nn,
                                     double* __restrict__ f, int outer) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nn * LM) return;
  const int64_t e = idx % nn;
  const int k = (int)(idx / nn);
  const int64_t dst = nd[2 * e], src = nd[2 * e + 1];
  if (outer) {
    const int64_t r = dst % ((int64_t)S * S);
    const int j = (int)(r / S) - H, i = (int)(r % S) - H;
    if (i >= 0 && i <= n && j >= 0 && j <= n) return;
  }
  f[(int64_t)k * NC + dst] = f[(int64_t)k * NC + src];
}

// step1 itype 1: gradient flux (q(nb) - q) / len * pedlen on every edge of
// the cells [-1, n]^2, the owner's lengths on halo edges
__global__ void dif_gradflux_kernel(DGeo g, int LM, const double* __restrict__ q, double* __restrict__ F) {
  int k, i, j;
  int64_t c;
  if (!square(g, -1, g.n + 2, LM, k, i, j, c)) return;
  const int64_t NC = g.NC, S = g.S;
  const double* qk = q + (int64_t)k * NC;
  double* Fk = F + (int64_t)k * 3 * NC;
  Fk[c] = (qk[c + 1] - qk[c]) / g.lenf[c] * g.pedf[c];
  Fk[NC + c] = (qk[c] - qk[c + S + 1]) / g.lenf[NC + c] * g.pedf[NC + c];
  Fk[2 * NC + c] = (qk[c + S] - qk[c]) / g.lenf[2 * NC + c] * g.pedf[2 * NC + c];
}

// halo edges := sign * owner's edge, per layer of an edge field [LM][3][NC]
__global__ void dif_edge_fill_kernel(int LM, int64_t NC, const int64_t* __restrict__ ed,
                                     const int8_t* __restrict__ es, int ne, double* __restrict__ F) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ne * LM) return;
  const int64_t e = idx % ne;
  const int k = (int)(idx / ne);
  double* Fk = F + (int64_t)k * 3 * NC;
  Fk[ed[2 * e]] = (double)es[e] * Fk[ed[2 * e + 1]];
}

// cal_div on the masters (owned cells and the two poles) of the node compute list
__global__ void dif_div_masters_kernel(DGeo g, int LM, const int64_t* __restrict__ ncl, int nmaster,
                                       const double* __restrict__ F, double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nmaster * LM) return;
  const int64_t c = ncl[2 * (idx % nmaster)];
  const int k = (int)(idx / nmaster);
  const int64_t NC = g.NC, S = g.S;
  const double* Fk = F + (int64_t)k * 3 * NC;
  out[(int64_t)k * NC + c] = (Fk[c] - Fk[c - 1] + Fk[NC + c - S - 1] - Fk[NC + c] + Fk[2 * NC + c] -
                              Fk[2 * NC + c - S]) *
                             g.hinv[c];
}

// vorticity of the edge field w on LORT (plane 0) and UPLT (plane 1) of the
// cells [-1, n]^2; the pentagon corners' merged triangles (post_step1 /
// post_vort) at the cells each corner sets
__global__ void dif_vort_kernel(DGeo g, int LM, const double* __restrict__ w, double* __restrict__ vort) {
  int k, i, j;
  int64_t c;
  if (!square(g, -1, g.n + 2, LM, k, i, j, c)) return;
  const int64_t NC = g.NC, S = g.S;
  const int n = g.n;
  const double* wk = w + (int64_t)k * 3 * NC;
  const double* tri = g.tri;
#define UV(e, x) (wk[(int64_t)(e) * NC + (x)] * g.len[(int64_t)(e) * NC + (x)])
  double vL, vU;
  if (i == -1 && j == -1) {  // IJMINUS: cell (0, 0)'s south-west triangle
    const int64_t idS = c + 1, idW = c + S;  // (0, -1), (-1, 0)
    vL = (UV(2, idS) - UV(0, idW) + UV(0, c)) / tri[c];
    vU = vL;
  } else {
    if (i == n - 1 && j == -1) {  // IPLUSJMINUS: id = (n, 0), idSW = (n - 1, -1)
      vL = (-UV(0, c + S + 1) + UV(0, c) + UV(1, c)) / tri[c];
    } else if (i == -1 && j == n) {  // IMINUSJPLUS: the value of UPLT(-1, n - 1)
      const int64_t id = c + 1, idSW = c - S;  // (0, n), (-1, n - 1)
      vL = (UV(2, id) - UV(1, idSW) - UV(2, idSW)) / tri[NC + idSW];
    } else if (i == n && j == n) {  // IJPLUS
      vL = (UV(0, c) - UV(0, c + S) - UV(2, c)) / tri[c];
    } else {
      vL = (UV(0, c) + UV(2, c + 1) + UV(1, c)) / tri[c];
    }
    if (i == n && j == -1) {  // IPLUSJMINUS: the value of LORT(n - 1, -1)
      const int64_t idSW = c - 1, id = c + S;  // (n - 1, -1), (n, 0)
      vU = (-UV(0, id) + UV(0, idSW) + UV(1, idSW)) / tri[idSW];
    } else if (i == -1 && j == n - 1) {  // IMINUSJPLUS: id = (0, n), idSW = (-1, n - 1)
      vU = (UV(2, c + S + 1) - UV(1, c) - UV(2, c)) / tri[NC + c];
    } else if (i == n && j == n) {  // IJPLUS
      vU = vL;
    } else {
      vU = -(UV(1, c) + UV(2, c) + UV(0, c + S)) / tri[NC + c];
    }
  }
#undef UV
  double* vk = vort + (int64_t)k * 2 * NC;
  vk[c] = vL;
  vk[NC + c] = vU;
}

// rot u = -(difference of the vorticity across the edge) / pedlen on the owned
// edges (cal_Laplacian_rotu, step1 itype 8); a zero dual length gives 0
__global__ void dif_rotu_kernel(DGeo g, int LM, const double* __restrict__ vort, double* __restrict__ rotu) {
  int k, i, j;
  int64_t c;
  if (!square(g, 0, g.n, LM, k, i, j, c)) return;
  const int64_t NC = g.NC, S = g.S;
  const double* vL = vort + (int64_t)k * 2 * NC;
  const double* vU = vL + NC;
  double* r = rotu + (int64_t)k * 3 * NC;
  const double p1 = g.ped[NC + c];
  r[c] = -(vL[c] - vU[c - S]) / g.ped[c];
  r[NC + c] = p1 != 0.0 ? -(vL[c] - vU[c]) / p1 : 0.0;
  r[2 * NC + c] = -(vL[c - 1] - vU[c]) / g.ped[2 * NC + c];
}

// the hooks' values: scalar diffusive fluxes on the edges of the cells
// [-1, n]^2 and the velocity source on the owned edges
__global__ void dif_phys_kernel(DGeo g, int LM, DPar p, const double* __restrict__ m, const double* __restrict__ t,
                                const double* __restrict__ lapm, const double* __restrict__ lapt,
                                const double* __restrict__ D, const double* __restrict__ vort,
                                double* __restrict__ phm, double* __restrict__ pht, double* __restrict__ phu) {
  int k, i, j;
  int64_t c;
  if (!square(g, -1, g.n + 2, LM, k, i, j, c)) return;
  const int64_t NC = g.NC, S = g.S;
  const int64_t nb[3] = {c + 1, c + S + 1, c + S};
  if (p.ls) {  // (-1)^Ls visc grad(q or Laplacian q) pedlen
    const double* qm = (p.ls == 1 ? m : lapm) + (int64_t)k * NC;
    const double* qt = (p.ls == 1 ? t : lapt) + (int64_t)k * NC;
    const double sg = (p.ls % 2) ? -1.0 : 1.0;
    for (int e = 0; e < 3; ++e) {
      const double dl = g.lenf[e * NC + c], pl = g.pedf[e * NC + c];
      const double gm = e == 1 ? (qm[c] - qm[nb[e]]) / dl : (qm[nb[e]] - qm[c]) / dl;
      const double gt = e == 1 ? (qt[c] - qt[nb[e]]) / dl : (qt[nb[e]] - qt[c]) / dl;
      phm[((int64_t)k * 3 + e) * NC + c] = sg * p.vm * gm * pl;
      pht[((int64_t)k * 3 + e) * NC + c] = sg * p.vt * gt * pl;
    }
  }
  if (p.lr && i >= 0 && i < g.n && j >= 0 && j < g.n) {  // (-1)^(Lr-1) (visc_divu grad D - visc_rotu curl W)
    double gd[3] = {0.0, 0.0, 0.0};
    if (p.ld) {
      const double* Dk = D + (int64_t)k * NC;
      gd[0] = (Dk[nb[0]] - Dk[c]) / g.len[c];
      gd[1] = (Dk[c] - Dk[nb[1]]) / g.len[NC + c];
      gd[2] = (Dk[nb[2]] - Dk[c]) / g.len[2 * NC + c];
    }
    const double* vL = vort + (int64_t)k * 2 * NC;
    const double* vU = vL + NC;
    const double cr[3] = {(vL[c] - vU[c - S]) / g.ped[c], (vL[c] - vU[c]) / g.ped[NC + c],
                          (vL[c - 1] - vU[c]) / g.ped[2 * NC + c]};
    const double sg = (p.lr % 2) ? 1.0 : -1.0;
    for (int e = 0; e < 3; ++e) phu[((int64_t)k * 3 + e) * NC + c] = sg * (p.vd * gd[e] - p.vr * cr[e]);
  }
}

DGeo dgeo(const wtg_ctx* c) {
  return DGeo{c->d_len, c->d_ped, c->d_lenf, c->d_pedf, c->d_hinv, c->d_tri, c->info.ncells, c->info.n, c->info.S};
}

int node_fill(wtg_ctx* c, int LM, double* f, int outer) {
  if (c->nn == 0) return 0;
  hipLaunchKernelGGL(dif_node_fill_kernel, dim3(nblocks((int64_t)c->nn * LM)), dim3(DNT), 0, c->stream, LM,
                     c->info.ncells, c->info.n, c->info.S, c->d_nd, c->nn, f, outer);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

int edge_fill(wtg_ctx* c, int LM, double* F) {
  if (c->ne == 0) return 0;
  hipLaunchKernelGGL(dif_edge_fill_kernel, dim3(nblocks((int64_t)c->ne * LM)), dim3(DNT), 0, c->stream, LM,
                     c->info.ncells, c->d_ed, c->d_es, c->ne, F);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

// div(grad q pedlen) of every layer of q into out, master values on the halo
int laplacian(wtg_ctx* c, int LM, const double* q, double* out) {
  const DGeo g = dgeo(c);
  const int64_t side = c->info.n + 2;
  hipLaunchKernelGGL(dif_gradflux_kernel, dim3(nblocks(10 * side * side * LM)), dim3(DNT), 0, c->stream, g, LM, q,
                     c->dif_flux);
  WTG_HIPCHK(c, hipGetLastError());
  if (int rc = edge_fill(c, LM, c->dif_flux)) return rc;
  hipLaunchKernelGGL(dif_div_masters_kernel, dim3(nblocks((int64_t)c->nmaster * LM)), dim3(DNT), 0, c->stream, g, LM,
                     c->d_ncl, c->nmaster, c->dif_flux, out);
  WTG_HIPCHK(c, hipGetLastError());
  return node_fill(c, LM, out, 0);
}

int vorticity(wtg_ctx* c, int LM, const double* w, double* vort) {
  const int64_t side = c->info.n + 2;
  hipLaunchKernelGGL(dif_vort_kernel, dim3(nblocks(10 * side * side * LM)), dim3(DNT), 0, c->stream, dgeo(c), LM, w,
                     vort);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}

}  // namespace

int wtg_diffusion_alloc(wtg_ctx* c) {
  const wtg_diffusion& d = c->dif;
  const int64_t NC = c->info.ncells;
  const size_t LM = (size_t)c->params.nlayers * c->nmembers;
  auto need = [&](double** p, size_t count) -> int {
    if (*p) return 0;
    if (int rc = wtg_dalloc<double>(c, p, count)) return rc;
    WTG_HIPCHK(c, hipMemsetAsync(*p, 0, count * sizeof(double), c->stream));
    return 0;
  };
  int rc = 0;
  if (d.laplace_divu && (rc = need(&c->dif_divu, LM * NC))) return rc;
  if (d.laplace_divu == 2 && (rc = need(&c->dif_lapd, LM * NC))) return rc;
  if (d.laplace_sclr == 2 && ((rc = need(&c->dif_lapm, LM * NC)) || (rc = need(&c->dif_lapt, LM * NC)))) return rc;
  if ((d.laplace_divu == 2 || d.laplace_sclr == 2) && (rc = need(&c->dif_flux, 3 * LM * NC))) return rc;
  if (d.laplace_rotu && (rc = need(&c->dif_vort, 2 * LM * NC))) return rc;
  if (d.laplace_rotu == 2 && (rc = need(&c->dif_rotu, 3 * LM * NC))) return rc;
  if (d.laplace_sclr && ((rc = need(&c->dif_phm, 3 * LM * NC)) || (rc = need(&c->dif_pht, 3 * LM * NC)))) return rc;
  if (d.laplace_rotu && (rc = need(&c->dif_phu, 3 * LM * NC))) return rc;
  WTG_HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int wtg_diffusion_prepass(wtg_ctx* c, const double* m, const double* t, const double* u) {
  const wtg_diffusion& d = c->dif;
  const int LM = c->params.nlayers * c->nmembers;
  const DGeo g = dgeo(c);
  const int64_t n = c->info.n;
  if (d.laplace_divu) {
    hipLaunchKernelGGL(dif_divu_kernel, dim3(nblocks(10 * (n + 1) * (n + 1) * LM)), dim3(DNT), 0, c->stream, g, LM,
                       u, c->dif_divu);
    WTG_HIPCHK(c, hipGetLastError());
    if (int rc = node_fill(c, LM, c->dif_divu, 1)) return rc;
    if (d.laplace_divu == 2)
      if (int rc = laplacian(c, LM, c->dif_divu, c->dif_lapd)) return rc;
  }
  if (d.laplace_sclr == 2) {
    if (int rc = laplacian(c, LM, m, c->dif_lapm)) return rc;
    if (int rc = laplacian(c, LM, t, c->dif_lapt)) return rc;
  }
  if (d.laplace_rotu) {
    if (int rc = vorticity(c, LM, u, c->dif_vort)) return rc;
    if (d.laplace_rotu == 2) {
      hipLaunchKernelGGL(dif_rotu_kernel, dim3(nblocks(10 * n * n * LM)), dim3(DNT), 0, c->stream, g, LM, c->dif_vort,
                         c->dif_rotu);
      WTG_HIPCHK(c, hipGetLastError());
      if (int rc = edge_fill(c, LM, c->dif_rotu)) return rc;
      if (int rc = vorticity(c, LM, c->dif_rotu, c->dif_vort)) return rc;
    }
  }
  const DPar p{d.laplace_sclr, d.laplace_divu, d.laplace_rotu, d.visc_mass, d.visc_temp, d.visc_divu, d.visc_rotu};
  hipLaunchKernelGGL(dif_phys_kernel, dim3(nblocks(10 * (n + 2) * (n + 2) * LM)), dim3(DNT), 0, c->stream, g, LM, p,
                     m, t, c->dif_lapm, c->dif_lapt, d.laplace_divu == 2 ? c->dif_lapd : c->dif_divu, c->dif_vort,
                     c->dif_phm, c->dif_pht, c->dif_phu);
  WTG_HIPCHK(c, hipGetLastError());
  return 0;
}
