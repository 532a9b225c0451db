// Multi-level explicit step on a full hierarchy of uniform levels (the
// reference's configuration with every mask TOLRNZ: a uniform refinement
// with zero thresholds, SURVEY config levels J5..J7 / J5..J8):
//
//   trend_ml over levels with compute-or-restrict   multi_level.f90:5-188, 311-577
//   forward_wavelet_transform                       wavelet.f90:20-101
//   inverse_wavelet_transform                       wavelet.f90:208-296
//   WT_after_step                                   adapt.f90:192-237
//   RK4                                             time_integr.f90:78-103
//
// One context per level (coarse to fine) holds the level's state slots.
// The finest level's trend is the layer-pair stage kernel, which also keeps
// the fluxes, Bernoulli and Exner functions, velocity source and scalar
// trends that the coarser levels restrict; on a full hierarchy every coarse
// value is a restriction (sub-sampled Bernoulli and Exner functions,
// conservative fluxes with the R_F_wgt correction, summed velocity
// sources), so the coarse levels never evaluate step1.  Every kernel is
// layer-batched (one launch per operator for all layers) and FP64 in the
// reference's evaluation order.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <vector>

#include "ctx.h"
#include "mg_ops.h"

struct wtg_ml {
  int nl = 0, L = 0;
  std::vector<wtg_ctx*> lv;
  std::vector<hipStream_t> saved_stream;
  // transfer data of the pair (l-1, l), stored with level l (rf with l-1)
  std::vector<double*> ovl, split, iu, rf;
  std::vector<int64_t*> wide;
  int WH = 0;
  // work per level
  std::vector<double*> hflux;  // [2L][3][NC]: mass fluxes of every layer, then temperature
  std::vector<double*> dsc;    // [2L][NC]: scalar trends (mass, then temperature)
  std::vector<double*> src;    // [L][3][NC]: velocity sources
  std::vector<double*> wcs;    // [2L][NC]: scalar wavelets (l >= 1)
  std::vector<double*> wcu;    // [L][3][NC]: velocity wavelets (l >= 1)
  double *bern = nullptr, *exn = nullptr;  // finest level [L][NC]
  std::vector<double*> eg;     // [NC] free surface of the gradient (coarse levels sub-sampled)
  std::vector<void*> allocs;
};

namespace {

constexpr int NT = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

// shared.f90:282-330 tables, 0-based edge index e = RT 0, DG 1, UP 2
__constant__ int c_HSO[3] = {2, 0, 4};
__constant__ int c_HEXS[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {-1, 0, 0}, {-1, -1, 1},
                                  {0, -1, 2}, {0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {-1, 0, 0}};
__constant__ int c_NGHB[10][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 0}};
__constant__ int c_ENDPT[3][2][2] = {{{0, 0}, {1, 0}}, {{1, 1}, {0, 0}}, {{0, 0}, {0, 1}}};
__constant__ int c_OPPNO[3][2][2] = {{{0, -1}, {1, 1}}, {{0, 1}, {1, 0}}, {{1, 1}, {-1, 0}}};
__constant__ int c_BFLY2[3][4][2] = {{{-3, -2}, {1, -2}, {3, 2}, {-1, 2}},
                                     {{1, 3}, {-3, -1}, {-1, -3}, {3, 1}},
                                     {{2, -1}, {2, 3}, {-2, 1}, {-2, -3}}};
// Iu_Base_Wgt (wavelet.f90:1765)
__constant__ double c_BASE[9] = {16.0 / 16, -1.0 / 16, 1.0 / 16, 1.0 / 16, -1.0 / 16,
                                 -1.0 / 16, -1.0 / 16, 1.0 / 16, 1.0 / 16};

// geometry of a level pair for the velocity and flux kernels
struct PairGeo {
  int nc, Sc, nf, Sf;
  int64_t NCc, NCf;
  const double *len_c, *tri_c;  // coarse [3][NCc], [2][NCc]
  const double *len_f, *tri_f;  // fine
  const double* iu;             // [9][NCf]
  const double *ovl, *split;    // [4][NCf], [2][NCf]
  const double* hinv_f;
  const double* rf;             // [24][NCc]
  const int64_t* wide;          // fine level's lattice map, frame [-WH, nf+WH)
  int WH;
};

__device__ __forceinline__ int64_t cellc(const PairGeo& g, int lz, int i, int j) {
  return (int64_t)lz * g.Sc * g.Sc + (int64_t)(j + 2) * g.Sc + (i + 2);
}
__device__ __forceinline__ int64_t cellf(const PairGeo& g, int lz, int i, int j) {
  return (int64_t)lz * g.Sf * g.Sf + (int64_t)(j + 2) * g.Sf + (i + 2);
}

// ---------------------------------------------------------------------------
// scalar transform, layer-batched: field-layers y = field * L + layer (field
// 0 mass, 1 temperature), ML_KC consecutive ones per thread (blockIdx.y =
// chunk), so a thread's list entry and stencil geometry serve all of them;
// compute-list threads evaluate at src, write dst
// ---------------------------------------------------------------------------
constexpr int ML_KC = 6;
__host__ __device__ constexpr int ml_chunks(int nfl) { return (nfl + ML_KC - 1) / ML_KC; }
// the velocity transform's gathers need the threads in flight: one layer
// per thread (two or six ran slower at J8 L30)
constexpr int ML_KCV = 1;

__global__ void ml_scalar_wavelet_kernel(int n, const int64_t* __restrict__ ncl, MgGeo g, int L,
                                         const double* __restrict__ sm, const double* __restrict__ st,
                                         double* __restrict__ wc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t x = ncl[2 * t + 1], dst = ncl[2 * t];
  int lz, i, j;
  mg_ij(x, g.Sf, lz, i, j);
  const bool mid = (i | j) & 1;
  const int y0 = blockIdx.y * ML_KC, y1 = min(y0 + ML_KC, 2 * L);
  for (int y = y0; y < y1; ++y) {
    const double* s = (y < L ? sm : st) + (int64_t)(y % L) * g.NCf;
    wc[(int64_t)y * g.NCf + dst] = mid ? s[x] - mg_interp(g, s, x, i, j) : 0.0;
  }
}

__global__ void ml_restrict_scalar_kernel(int n, const int64_t* __restrict__ ncl, MgGeo g, int L,
                                          const double* __restrict__ fm, const double* __restrict__ ft,
                                          const double* __restrict__ wc, double* __restrict__ cm,
                                          double* __restrict__ ct, int64_t NCc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t p = ncl[2 * t + 1], dst = ncl[2 * t];
  int lz, i, j;
  mg_ij(p, g.Sc, lz, i, j);
  const int64_t c = (int64_t)lz * g.Sf * g.Sf + (int64_t)(2 * j + 2) * g.Sf + (2 * i + 2);
  const bool pole = p == g.sp_c || p == g.np_c;
  const int y0 = blockIdx.y * ML_KC, y1 = min(y0 + ML_KC, 2 * L);
  for (int y = y0; y < y1; ++y) {
    const int k = y % L;
    const double* sf = (y < L ? fm : ft) + (int64_t)k * g.NCf;
    double* sc = (y < L ? cm : ct) + (int64_t)k * NCc;
    const double* w = wc + (int64_t)y * g.NCf;
    const double lift = pole ? mg_pole_lift(g, w, p == g.sp_c, false) : mg_lift(g, w, c, true);
    sc[dst] = sf[c] + lift * g.hinv_c[p];
  }
}

__global__ void ml_prolong_even_kernel(int n, const int64_t* __restrict__ ncl, MgGeo g, int L,
                                       const double* __restrict__ cm, const double* __restrict__ ct, int64_t NCc,
                                       const double* __restrict__ wc, double* __restrict__ fm,
                                       double* __restrict__ ft) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t x = ncl[2 * t + 1], dst = ncl[2 * t];
  int lz, i, j;
  mg_ij(x, g.Sf, lz, i, j);
  if ((i | j) & 1) return;
  const int64_t p = (int64_t)lz * g.Sc * g.Sc + (int64_t)(j / 2 + 2) * g.Sc + (i / 2 + 2);
  const bool pole = x == g.sp_f || x == g.np_f;
  const int y0 = blockIdx.y * ML_KC, y1 = min(y0 + ML_KC, 2 * L);
  for (int y = y0; y < y1; ++y) {
    const int k = y % L;
    const double* sc = (y < L ? cm : ct) + (int64_t)k * NCc;
    double* sf = (y < L ? fm : ft) + (int64_t)k * g.NCf;
    const double* w = wc + (int64_t)y * g.NCf;
    const double lift = pole ? mg_pole_lift(g, w, x == g.sp_f, true) : mg_lift(g, w, x, false);
    sf[dst] = sc[p] - lift * g.hinv_c[p];
  }
}

__global__ void ml_prolong_odd_kernel(int n, const int64_t* __restrict__ ncl, MgGeo g, int L,
                                      const double* __restrict__ wc, double* __restrict__ fm,
                                      double* __restrict__ ft) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t x = ncl[2 * t + 1], dst = ncl[2 * t];
  int lz, i, j;
  mg_ij(x, g.Sf, lz, i, j);
  if (!((i | j) & 1)) return;
  const int y0 = blockIdx.y * ML_KC, y1 = min(y0 + ML_KC, 2 * L);
  for (int y = y0; y < y1; ++y) {
    double* sf = (y < L ? fm : ft) + (int64_t)(y % L) * g.NCf;
    sf[dst] = mg_interp(g, sf, x, i, j) + wc[(int64_t)y * g.NCf + x];
  }
}

// ---------------------------------------------------------------------------
// velocity transform: one thread per (coarse owned node, layer)
// ---------------------------------------------------------------------------
// Interp_outer_velo (wavelet.f90:1054-1094)
__device__ __forceinline__ double interp_outer(const PairGeo& g, const double* __restrict__ uc, int lz, int ip,
                                               int jp, int e, int64_t id2) {
  const int so = c_HSO[e];
  const int* e1 = c_ENDPT[e][0];
  const int* e2 = c_ENDPT[e][1];
  const int* o1 = c_OPPNO[e][0];
  const int* o2 = c_OPPNO[e][1];
  auto vc = [&](const int* o, int col) {
    return uc[(int64_t)c_HEXS[col][2] * g.NCc + cellc(g, lz, ip + o[0] + c_HEXS[col][0], jp + o[1] + c_HEXS[col][1])];
  };
  double v[9];
  v[0] = uc[(int64_t)e * g.NCc + cellc(g, lz, ip, jp)];
  v[1] = vc(e2, so + 2);
  v[2] = vc(e1, so + 3);
  v[3] = vc(e1, so + 5);
  v[4] = vc(e2, so + 0);
  v[5] = vc(o1, so + 1) - vc(e1, so + 2);
  v[6] = vc(e2, so + 3) - vc(o1, so + 4);
  v[7] = vc(o2, so + 4) - vc(e2, so + 5);
  v[8] = vc(e1, so + 0) - vc(o2, so + 1);
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    const double w = c_BASE[m] + g.iu[(int64_t)m * g.NCf + id2];
    s = s + w * v[m];
  }
  return s;
}

// Interp_inner_velo (wavelet.f90:1000-1052); slot[] = the six inner child edges
__device__ __forceinline__ void interp_inner(const PairGeo& g, const double* __restrict__ uc,
                                             const double* __restrict__ uf, int lz, int ip, int jp, double out[6],
                                             int64_t slot[6]) {
  const int64_t idp = cellc(g, lz, ip, jp);
  double curl[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int64_t id1 = cellc(g, lz, ip - t + 1, jp), id2 = cellc(g, lz, ip, jp + t);
    curl[t] = (uc[g.NCc + idp] * g.len_c[g.NCc + idp] + uc[2 * g.NCc + id1] * g.len_c[2 * g.NCc + id1] +
               uc[id2] * g.len_c[id2]) /
              g.tri_c[(int64_t)t * g.NCc + idp];
  }
  const int ic = 2 * ip, jc = 2 * jp;
  const int64_t NF = g.NCf;
  const int64_t id = cellf(g, lz, ic, jc), idE = cellf(g, lz, ic + 1, jc), idNE = cellf(g, lz, ic + 1, jc + 1),
                idN = cellf(g, lz, ic, jc + 1), idN2E = cellf(g, lz, ic + 2, jc + 1),
                id2NE = cellf(g, lz, ic + 1, jc + 2), idN2 = cellf(g, lz, ic, jc + 2), idE2 = cellf(g, lz, ic + 2, jc);
  const double* len = g.len_f;
  const double* tri = g.tri_f;
#define UF(e, c) (uf[(int64_t)(e) * NF + (c)])
#define LF(e, c) (len[(int64_t)(e) * NF + (c)])
  out[0] = (tri[id] * curl[0] - UF(0, id) * LF(0, id) - UF(1, id) * LF(1, id)) / LF(2, idE);
  out[1] = (tri[idE] * curl[0] - UF(0, idE) * LF(0, idE) - UF(2, idE2) * LF(2, idE2)) / LF(1, idE);
  out[2] = (tri[idNE] * curl[0] - UF(1, idNE) * LF(1, idNE) - UF(2, idN2E) * LF(2, idN2E)) / LF(0, idNE);
  out[3] = (tri[NF + id] * curl[1] - UF(2, id) * LF(2, id) - UF(1, id) * LF(1, id)) / LF(0, idN);
  out[4] = (tri[NF + idN] * curl[1] - UF(2, idN) * LF(2, idN) - UF(0, idN2) * LF(0, idN2)) / LF(1, idN);
  out[5] = (tri[NF + idNE] * curl[1] - UF(1, idNE) * LF(1, idNE) - UF(0, id2NE) * LF(0, id2NE)) / LF(2, idNE);
#undef UF
#undef LF
  slot[0] = 2 * NF + idE;
  slot[1] = NF + idE;
  slot[2] = idNE;
  slot[3] = idN;
  slot[4] = NF + idN;
  slot[5] = 2 * NF + idNE;
}

__device__ __forceinline__ bool parent_of(const PairGeo& g, int& lz, int& ip, int& jp, int& k, int L) {
  const int64_t per = (int64_t)10 * g.nc * g.nc;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per) return false;
  k = blockIdx.y;
  lz = (int)(t / ((int64_t)g.nc * g.nc));
  const int r = (int)(t % ((int64_t)g.nc * g.nc));
  jp = r / g.nc;
  ip = r % g.nc;
  (void)L;
  return true;
}

// The velocity transform kernels: a 256-thread block is 64 consecutive
// parents x 4 layers, one wave per layer (blockIdx.y: group of 4 layers), so
// the state planes are read along i by every wave and the layer-invariant
// geometry a parent needs (I_u_wgt, lengths, areas: ~420 bytes) is fetched
// once from L2 for the four layers (the L1 serves the other three waves)
// instead of once per layer.
constexpr int VPB = 64;         // parents per block
constexpr int VLB = NT / VPB;   // layers per block (one wave each)
inline dim3 velo_grid(int nc, int L) {
  return dim3((unsigned)(((int64_t)10 * nc * nc + VPB - 1) / VPB), (unsigned)((L + VLB - 1) / VLB));
}
__device__ __forceinline__ bool parent_layer_of(const PairGeo& g, int& lz, int& ip, int& jp, int& k, int L) {
  const int64_t per = (int64_t)10 * g.nc * g.nc;
  const int64_t t = (int64_t)blockIdx.x * VPB + threadIdx.x % VPB;
  k = blockIdx.y * VLB + threadIdx.x / VPB;
  if (t >= per || k >= L) return false;
  lz = (int)(t / ((int64_t)g.nc * g.nc));
  const int r = (int)(t % ((int64_t)g.nc * g.nc));
  jp = r / g.nc;
  ip = r % g.nc;
  return true;
}
static_assert(ML_KCV == 1, "the velocity transform kernels take one layer per thread");

// Compute_velo_wavelets (wavelet.f90:672-707): 6 outer and 6 inner child edges
__global__ void ml_velo_wavelet_kernel(PairGeo g, int L, const double* __restrict__ ucl,
                                       const double* __restrict__ ufl, double* __restrict__ wcl) {
  int lz, ip, jp, chunk;
  if (!parent_layer_of(g, lz, ip, jp, chunk, L)) return;
  const int ic = 2 * ip, jc = 2 * jp;
  const int k1 = min(chunk * ML_KCV + ML_KCV, L);
  for (int k = chunk * ML_KCV; k < k1; ++k) {
    const double* uc = ucl + (int64_t)k * 3 * g.NCc;
    const double* uf = ufl + (int64_t)k * 3 * g.NCf;
    double* wc = wcl + (int64_t)k * 3 * g.NCf;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const int64_t id1 = cellf(g, lz, ic + c_ENDPT[e][0][0], jc + c_ENDPT[e][0][1]);
      const int64_t id2 = cellf(g, lz, ic + c_ENDPT[e][1][0], jc + c_ENDPT[e][1][1]);
      const double u = interp_outer(g, uc, lz, ip, jp, e, id2);
      const double w2 = uf[(int64_t)e * g.NCf + id2] - u;
      wc[(int64_t)e * g.NCf + id2] = w2;
      wc[(int64_t)e * g.NCf + id1] = -w2;
    }
    double ui[6];
    int64_t sl[6];
    interp_inner(g, uc, uf, lz, ip, jp, ui, sl);
#pragma unroll
    for (int q = 0; q < 6; ++q) wc[sl[q]] = uf[sl[q]] - ui[q];
  }
}

// Reconstruct_outer_velo (wavelet.f90:954-976)
__global__ void ml_velo_outer_kernel(PairGeo g, int L, const double* __restrict__ ucl,
                                     const double* __restrict__ wcl, double* __restrict__ ufl) {
  int lz, ip, jp, chunk;
  if (!parent_layer_of(g, lz, ip, jp, chunk, L)) return;
  const int ic = 2 * ip, jc = 2 * jp;
  const int64_t idp = cellc(g, lz, ip, jp);
  const int k1 = min(chunk * ML_KCV + ML_KCV, L);
  for (int k = chunk * ML_KCV; k < k1; ++k) {
    const double* uc = ucl + (int64_t)k * 3 * g.NCc;
    const double* wc = wcl + (int64_t)k * 3 * g.NCf;
    double* uf = ufl + (int64_t)k * 3 * g.NCf;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const int64_t id1 = cellf(g, lz, ic + c_ENDPT[e][0][0], jc + c_ENDPT[e][0][1]);
      const int64_t id2 = cellf(g, lz, ic + c_ENDPT[e][1][0], jc + c_ENDPT[e][1][1]);
      const double u = interp_outer(g, uc, lz, ip, jp, e, id2);
      const double v2 = u + wc[(int64_t)e * g.NCf + id2];
      uf[(int64_t)e * g.NCf + id2] = v2;
      uf[(int64_t)e * g.NCf + id1] = 2.0 * uc[(int64_t)e * g.NCc + idp] - v2;
    }
  }
}

// Reconstruct_inner_velo (wavelet.f90:978-998); outer edges ghost filled
__global__ void ml_velo_inner_kernel(PairGeo g, int L, const double* __restrict__ ucl,
                                     const double* __restrict__ wcl, double* __restrict__ ufl) {
  int lz, ip, jp, chunk;
  if (!parent_layer_of(g, lz, ip, jp, chunk, L)) return;
  const int k1 = min(chunk * ML_KCV + ML_KCV, L);
  for (int k = chunk * ML_KCV; k < k1; ++k) {
    const double* uc = ucl + (int64_t)k * 3 * g.NCc;
    const double* wc = wcl + (int64_t)k * 3 * g.NCf;
    double* uf = ufl + (int64_t)k * 3 * g.NCf;
    double ui[6];
    int64_t sl[6];
    interp_inner(g, uc, uf, lz, ip, jp, ui, sl);
#pragma unroll
    for (int q = 0; q < 6; ++q) uf[sl[q]] = ui[q] + wc[sl[q]];
  }
}

// pentagon corrections at the three lozenge corners that have them:
// Compute_velo_wavelets_penta (sign +1, on the wavelets) and
// Reconstruct_velo_penta (sign -1, on the velocities), wavelet.f90:709-781,
// 897-925, 1096-1152.  One thread per (lozenge, layer).
__global__ void ml_velo_penta_kernel(PairGeo g, int L, int sign, const double* __restrict__ ucl,
                                     double* __restrict__ fl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 10 * L) return;
  const int lz = t % 10, k = t / 10;
  const double* uc = ucl + (int64_t)k * 3 * g.NCc;
  double* f = fl + (int64_t)k * 3 * g.NCf;
  const int nc = g.nc, nf = g.nf;
  const int64_t NF = g.NCf;
#define UC(e, i, j) uc[(int64_t)(e) * g.NCc + cellc(g, lz, i, j)]
  {  // upper-left corner (IMINUSJPLUS)
    const int64_t id = cellf(g, lz, 0, nf - 2), idN = cellf(g, lz, 0, nf - 1);
    double v = (c_BASE[7] + g.iu[7 * NF + idN]) * (UC(2, 0, nc) + UC(0, -1, nc));
    if (sign > 0) v = -v;
    f[2 * NF + id] = f[2 * NF + id] - v;
    f[2 * NF + idN] = f[2 * NF + idN] + v;
  }
  {  // lower-right corner (IPLUSJMINUS)
    const int64_t id = cellf(g, lz, nf - 2, 0), idE = cellf(g, lz, nf - 1, 0);
    double v = (c_BASE[6] + g.iu[6 * NF + idE]) * (UC(0, nc, 0) + UC(2, nc, -1));
    if (sign < 0) v = -v;
    f[id] = f[id] - v;
    f[idE] = f[idE] + v;
  }
  {  // lower-left corner (IJMINUS): velo_interp_penta_corr
    const int64_t id = cellf(g, lz, 0, 0), idN = cellf(g, lz, 0, 1), idE = cellf(g, lz, 1, 0);
    const double u1 =
        (c_BASE[8] + g.iu[8 * NF + idN]) * ((-UC(2, 0, -1) - (-UC(0, -1, -1))) - (UC(1, -1, -1) - UC(2, -1, -1)));
    const double u2 = (c_BASE[5] + g.iu[5 * NF + idE]) * (UC(0, -1, -1) + UC(0, -1, 0) - (UC(0, -1, -1) - UC(1, -1, -1)));
    if (sign > 0) {
      f[2 * NF + id] = f[2 * NF + id] + u1;
      f[2 * NF + idN] = f[2 * NF + idN] - u1;
      f[id] = f[id] + u2;
      f[idE] = f[idE] - u2;
    } else {
      f[2 * NF + id] = f[2 * NF + id] - u1;
      f[2 * NF + idN] = f[2 * NF + idN] + u1;
      f[id] = f[id] - u2;
      f[idE] = f[idE] + u2;
    }
  }
#undef UC
}

// ---------------------------------------------------------------------------
// flux restriction (flux_cpt_restr, multi_level.f90:421-577): one thread per
// (coarse owned node, field-layer); blockIdx.y = field * L + layer
// ---------------------------------------------------------------------------
enum { UZM = 0, UPZ, UMZ, UZP, VMM, VPM, VMP, VPP, WMM, WPM, WMP, WPP, WMMM, WPPP, VMPP, VPMM };

struct Ind {
  int64_t ed[16];
  int64_t mp, pp, pm, mm;
};

// get_indices (multi_level.f90:619-659) at fine (i, j), edge e
__device__ __forceinline__ void get_indices(const PairGeo& g, int lz, int i, int j, int e, Ind& x) {
  const int so = c_HSO[e];
  const int64_t NF = g.NCf;
  auto ed = [&](int bi, int bj, int col) {
    return (int64_t)c_HEXS[col][2] * NF + cellf(g, lz, bi + c_HEXS[col][0], bj + c_HEXS[col][1]);
  };
  x.ed[UMZ] = ed(i, j, so + 1);
  x.ed[UPZ] = ed(i, j, so + 4);
  x.ed[WMP] = ed(i, j, so + 0);
  x.ed[VPP] = ed(i, j, so + 5);
  x.ed[WPM] = ed(i, j, so + 3);
  x.ed[VMM] = ed(i, j, so + 2);
  const int mpi = i + c_NGHB[so + 0][0], mpj = j + c_NGHB[so + 0][1];
  const int ppi = i + c_NGHB[so + 5][0], ppj = j + c_NGHB[so + 5][1];
  const int pmi = i + c_NGHB[so + 3][0], pmj = j + c_NGHB[so + 3][1];
  const int mmi = i + c_NGHB[so + 2][0], mmj = j + c_NGHB[so + 2][1];
  x.mp = cellf(g, lz, mpi, mpj);
  x.pp = cellf(g, lz, ppi, ppj);
  x.pm = cellf(g, lz, pmi, pmj);
  x.mm = cellf(g, lz, mmi, mmj);
  x.ed[VMP] = ed(mpi, mpj, so + 2);
  x.ed[VMPP] = ed(mpi, mpj, so + 5);
  x.ed[UZP] = ed(mpi, mpj, so + 4);
  x.ed[WPPP] = ed(ppi, ppj, so + 0);
  x.ed[WPP] = ed(ppi, ppj, so + 3);
  x.ed[VPM] = ed(pmi, pmj, so + 5);
  x.ed[VPMM] = ed(pmi, pmj, so + 2);
  x.ed[UZM] = ed(pmi, pmj, so + 1);
  x.ed[WMMM] = ed(mmi, mmj, so + 3);
  x.ed[WMM] = ed(mmi, mmj, so + 0);
}

// scalar trend at a fine lattice point, through the wide map beyond the tile
__device__ __forceinline__ double dval(const PairGeo& g, const double* __restrict__ df, int lz, int i, int j) {
  if (i >= -2 && j >= -2 && i < g.nf + 2 && j < g.nf + 2) return df[cellf(g, lz, i, j)];
  const int WS = g.nf + 2 * g.WH;
  const int64_t s = g.wide[(int64_t)lz * WS * WS + (int64_t)(j + g.WH) * WS + (i + g.WH)];
  return s >= 0 ? df[s] : 0.0;
}

__device__ __forceinline__ double part_coarse_flux(const PairGeo& g, const double* __restrict__ hf,
                                                   const double* __restrict__ df, int lz, int i, int j, int e) {
  const int64_t NF = g.NCf;
  const double* a = g.ovl;
  const double* sp = g.split;
  Ind x;
  get_indices(g, lz, i, j, e, x);
  const int64_t c = cellf(g, lz, i, j);
  double area0 = a[c], area1 = a[NF + c];
  double ol0 = sp[c], ol1 = sp[NF + c];
  double ol2 = a[2 * NF + c] - ol0, ol3 = a[3 * NF + c] - ol1;
  area0 = area0 + ol0 + ol3;
  area1 = area1 + ol1 + ol2;
  const double tot = area0 + area1;
  area0 = area0 / tot;
  area1 = area1 / tot;
  ol0 = sp[x.pp];
  ol1 = sp[NF + x.mm];
  ol2 = a[2 * NF + x.mp] - sp[x.mp];
  ol3 = a[3 * NF + x.pm] - sp[NF + x.pm];
  const double s1 = hf[x.ed[UPZ]] * area0 + hf[x.ed[UMZ]] * area1;
  const double s2 = hf[x.ed[VMM]] + hf[x.ed[WMP]];
  const double s3 = hf[x.ed[WPM]] + hf[x.ed[VPP]];
  return s1 - s2 * area1 - s3 * area0 + ol2 * df[x.mp] - ol3 * df[x.pm] - ol0 * df[x.pp] + ol1 * df[x.mm];
}

__device__ __forceinline__ double coarse_flux(const PairGeo& g, const double* __restrict__ df, int lz, int i, int j,
                                              int e) {
  const int64_t NF = g.NCf;
  const double* a = g.ovl;
  const double* hi = g.hinv_f;
  const int so = c_HSO[e];
  auto nb = [&](int k) { return cellf(g, lz, i + c_NGHB[k][0], j + c_NGHB[k][1]); };
  const int64_t mz = nb(so + 1), pz = nb(so + 4), mp = nb(so + 0), pp = nb(so + 5), pm = nb(so + 3), mm = nb(so + 2);
  const double dmm2 = dval(g, df, lz, i + c_BFLY2[e][0][0], j + c_BFLY2[e][0][1]);
  const double dpm2 = dval(g, df, lz, i + c_BFLY2[e][1][0], j + c_BFLY2[e][1][1]);
  const double dpp2 = dval(g, df, lz, i + c_BFLY2[e][2][0], j + c_BFLY2[e][2][1]);
  const double dmp2 = dval(g, df, lz, i + c_BFLY2[e][3][0], j + c_BFLY2[e][3][1]);
  const int64_t id = cellf(g, lz, i, j);
#define A(k, c) a[(int64_t)((k) - 1) * NF + (c)]
  return (A(1, id) * A(2, id) * hi[id] + A(2, mp) * A(3, mp) * hi[mp] + A(1, pp) * A(3, pp) * hi[pp] +
          A(1, pm) * A(4, pm) * hi[pm] + A(2, mm) * A(4, mm) * hi[mm]) *
             (df[pz] - df[mz]) +
         A(3, pp) * A(4, pp) * hi[pp] * 0.5 * (dpp2 - df[mz]) + A(3, pm) * A(4, pm) * hi[pm] * 0.5 * (dpm2 - df[mz]) +
         A(3, mp) * A(4, mp) * hi[mp] * 0.5 * (df[pz] - dmp2) + A(3, mm) * A(4, mm) * hi[mm] * 0.5 * (df[pz] - dmm2);
#undef A
}

__device__ __forceinline__ double sum3(const double* __restrict__ hf, const Ind& x, int a0, int a1, int a2,
                                       const double* __restrict__ rf, int64_t NCc, int64_t p) {
  double s = 0.0;
  s = s + hf[x.ed[a0]] * rf[p];
  s = s + hf[x.ed[a1]] * rf[NCc + p];
  s = s + hf[x.ed[a2]] * rf[2 * NCc + p];
  return s;
}
__device__ __forceinline__ double sum3d(const double* __restrict__ hf, const Ind& x, int a0, int a1, int a2, int b0,
                                        int b1, int b2, const double* __restrict__ rf, int64_t NCc, int64_t p) {
  double s = 0.0;
  s = s + (hf[x.ed[a0]] - hf[x.ed[b0]]) * rf[p];
  s = s + (hf[x.ed[a1]] - hf[x.ed[b1]]) * rf[NCc + p];
  s = s + (hf[x.ed[a2]] - hf[x.ed[b2]]) * rf[2 * NCc + p];
  return s;
}

// The restriction of one coarse node reads the same fine positions and
// geometry for every field and layer: those are resolved once per thread
// (fine-plane offsets as 32-bit indices, the area and overlap weights of
// part_coarse_flux and coarse_flux evaluated in the same order as above),
// then the thread runs over FR_KC consecutive field-layers (blockIdx.y =
// chunk), each a sequence of loads and the reference's sums.
constexpr int FR_KC = 6;

struct FluxEdgeW {  // part_coarse_flux and coarse_flux weights of one fine edge
  double area0, area1, ol0, ol1, ol2, ol3;  // normalised areas, overlap parts
  double c5, wpp, wpm, wmp, wmm;            // coarse_flux: sum of 5 products, half-weights
};

__device__ __forceinline__ void flux_edge_prep(const PairGeo& g, int lz, int i, int j, int e, const Ind& x,
                                               FluxEdgeW& w, int& mz, int& pz, int b2[4]) {
  const int64_t NF = g.NCf;
  const double* a = g.ovl;
  const double* sp = g.split;
  const double* hi = g.hinv_f;
  const int64_t c = cellf(g, lz, i, j);
  double area0 = a[c], area1 = a[NF + c];
  double ol0 = sp[c], ol1 = sp[NF + c];
  double ol2 = a[2 * NF + c] - ol0, ol3 = a[3 * NF + c] - ol1;
  area0 = area0 + ol0 + ol3;
  area1 = area1 + ol1 + ol2;
  const double tot = area0 + area1;
  w.area0 = area0 / tot;
  w.area1 = area1 / tot;
  w.ol0 = sp[x.pp];
  w.ol1 = sp[NF + x.mm];
  w.ol2 = a[2 * NF + x.mp] - sp[x.mp];
  w.ol3 = a[3 * NF + x.pm] - sp[NF + x.pm];
  const int so = c_HSO[e];
  mz = (int)cellf(g, lz, i + c_NGHB[so + 1][0], j + c_NGHB[so + 1][1]);
  pz = (int)cellf(g, lz, i + c_NGHB[so + 4][0], j + c_NGHB[so + 4][1]);
  const int64_t id = c, mp = x.mp, pp = x.pp, pm = x.pm, mm = x.mm;
#define A(k, cc) a[(int64_t)((k) - 1) * NF + (cc)]
  w.c5 = A(1, id) * A(2, id) * hi[id] + A(2, mp) * A(3, mp) * hi[mp] + A(1, pp) * A(3, pp) * hi[pp] +
         A(1, pm) * A(4, pm) * hi[pm] + A(2, mm) * A(4, mm) * hi[mm];
  w.wpp = A(3, pp) * A(4, pp) * hi[pp] * 0.5;
  w.wpm = A(3, pm) * A(4, pm) * hi[pm] * 0.5;
  w.wmp = A(3, mp) * A(4, mp) * hi[mp] * 0.5;
  w.wmm = A(3, mm) * A(4, mm) * hi[mm] * 0.5;
#undef A
  // the butterfly's second ring (dval): the tile, else the wide map, else 0 (-1)
  for (int q = 0; q < 4; ++q) {
    const int bi = i + c_BFLY2[e][q][0], bj = j + c_BFLY2[e][q][1];
    if (bi >= -2 && bj >= -2 && bi < g.nf + 2 && bj < g.nf + 2) {
      b2[q] = (int)cellf(g, lz, bi, bj);
    } else {
      const int WS = g.nf + 2 * g.WH;
      b2[q] = (int)g.wide[(int64_t)lz * WS * WS + (int64_t)(bj + g.WH) * WS + (bi + g.WH)];
    }
  }
}

// loads at 32-bit byte offsets from a wave-uniform base (one VGPR per index)
__device__ __forceinline__ double ldb(const double* p, uint32_t off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) + off);
}
constexpr uint32_t NO_CELL = 0xffffffffu;

// part_coarse_flux + coarse_flux of one fine edge on one field-layer
// (offsets in bytes; ed: UPZ, UMZ, VMM, WMP, WPM, VPP; cell: mp, pm, pp, mm)
__device__ __forceinline__ double flux_edge_apply(const double* hf, const double* df, const FluxEdgeW& w,
                                                  const uint32_t ed[6], const uint32_t cell[4], uint32_t mz,
                                                  uint32_t pz, const uint32_t b2[4]) {
  const double s1 = ldb(hf, ed[0]) * w.area0 + ldb(hf, ed[1]) * w.area1;
  const double s2 = ldb(hf, ed[2]) + ldb(hf, ed[3]);
  const double s3 = ldb(hf, ed[4]) + ldb(hf, ed[5]);
  const double pf = s1 - s2 * w.area1 - s3 * w.area0 + w.ol2 * ldb(df, cell[0]) - w.ol3 * ldb(df, cell[1]) -
                    w.ol0 * ldb(df, cell[2]) + w.ol1 * ldb(df, cell[3]);
  const double dmm2 = b2[0] != NO_CELL ? ldb(df, b2[0]) : 0.0;
  const double dpm2 = b2[1] != NO_CELL ? ldb(df, b2[1]) : 0.0;
  const double dpp2 = b2[2] != NO_CELL ? ldb(df, b2[2]) : 0.0;
  const double dmp2 = b2[3] != NO_CELL ? ldb(df, b2[3]) : 0.0;
  const double dz = ldb(df, mz), dp = ldb(df, pz);
  const double cf = w.c5 * (dp - dz) + w.wpp * (dpp2 - dz) + w.wpm * (dpm2 - dz) + w.wmp * (dp - dmp2) +
                    w.wmm * (dp - dmm2);
  return pf + cf;
}

// sum3 / sum3d of R_F_wgt set q (rf planes 3q..3q+2 at byte offset po)
__device__ __forceinline__ double sum3i(const double* hf, const uint32_t a[3], const double* rf, int64_t NCc,
                                        uint32_t po) {
  double s = 0.0;
  s = s + ldb(hf, a[0]) * ldb(rf, po);
  s = s + ldb(hf, a[1]) * ldb(rf + NCc, po);
  s = s + ldb(hf, a[2]) * ldb(rf + 2 * NCc, po);
  return s;
}
__device__ __forceinline__ double sum3di(const double* hf, const uint32_t a[3], const uint32_t b[3],
                                         const double* rf, int64_t NCc, uint32_t po) {
  double s = 0.0;
  s = s + (ldb(hf, a[0]) - ldb(hf, b[0])) * ldb(rf, po);
  s = s + (ldb(hf, a[1]) - ldb(hf, b[1])) * ldb(rf + NCc, po);
  s = s + (ldb(hf, a[2]) - ldb(hf, b[2])) * ldb(rf + 2 * NCc, po);
  return s;
}

// sum3 + sum3d of one R_F_wgt pair (sets q, q + 1) at fine index set x:
// sign * sum3(a) + sign * sum3d(c, d) in the reference's order
struct SmOps {
  uint32_t a[3], c[3], d[3];
};

// the restriction of coarse edge E of one coarse node (blockIdx.z = E): its
// part_coarse_flux + coarse_flux and the two sm terms it shares with the
// neighbouring edges, hc(E) = ((pf + cf) + sm_E) + sm_{E+1}
template <int E>
__device__ __forceinline__ void flux_restrict_edge(const PairGeo& g, int nfl, const double* __restrict__ hfl,
                                                   const double* __restrict__ dfl, double* __restrict__ hcl, int lz,
                                                   int ip, int jp, int chunk) {
  const int64_t NCc = g.NCc;
  const int i = 2 * ip, j = 2 * jp;
  const int64_t p = cellc(g, lz, ip, jp);
  auto B = [](int64_t k) { return (uint32_t)(k * 8); };
  // sm_E, sm_{E+1}: sm0, sm1 from the RT child (i+1, j), sm2, sm3 from the UP child (i, j+1)
  SmOps sa, sb;
  int qa, qb;
  double sga, sgb;
  {
    Ind x0, x2;
    if (E <= 1) get_indices(g, lz, i + 1, j, 0, x0);
    if (E >= 1) get_indices(g, lz, i, j + 1, 2, x2);
    auto set = [&](SmOps& o, int k) {
      if (k == 0) o = SmOps{{B(x0.ed[WPM]), B(x0.ed[UZM]), B(x0.ed[VMM])},
                            {B(x0.ed[VPM]), B(x0.ed[WMMM]), B(x0.ed[UMZ])},
                            {B(x0.ed[UPZ]), B(x0.ed[VPMM]), B(x0.ed[WMM])}};
      if (k == 1) o = SmOps{{B(x0.ed[WMP]), B(x0.ed[UZP]), B(x0.ed[VPP])},
                            {B(x0.ed[VMP]), B(x0.ed[WPPP]), B(x0.ed[UPZ])},
                            {B(x0.ed[UMZ]), B(x0.ed[VMPP]), B(x0.ed[WPP])}};
      if (k == 2) o = SmOps{{B(x2.ed[UZM]), B(x2.ed[VMM]), B(x2.ed[WPM])},
                            {B(x2.ed[WMMM]), B(x2.ed[UMZ]), B(x2.ed[VPM])},
                            {B(x2.ed[VPMM]), B(x2.ed[WMM]), B(x2.ed[UPZ])}};
      if (k == 3) o = SmOps{{B(x2.ed[UZP]), B(x2.ed[VPP]), B(x2.ed[WMP])},
                            {B(x2.ed[WPPP]), B(x2.ed[UPZ]), B(x2.ed[VMP])},
                            {B(x2.ed[VMPP]), B(x2.ed[WPP]), B(x2.ed[UMZ])}};
    };
    set(sa, E);
    set(sb, E + 1);
    qa = 6 * E, qb = 6 * (E + 1);
    sga = (E == 0 || E == 2) ? -1.0 : 1.0;  // sm0 and sm2 enter with a minus sign
    sgb = (E + 1 == 2) ? -1.0 : 1.0;
  }
  FluxEdgeW w;
  uint32_t ed[6], cl[4], mz, pz, b2[4];
  {
    const int fi = E == 2 ? i : i + 1, fj = E == 0 ? j : j + 1;
    Ind x;
    get_indices(g, lz, fi, fj, E, x);
    int m_, p_, bb[4];
    flux_edge_prep(g, lz, fi, fj, E, x, w, m_, p_, bb);
    mz = B(m_), pz = B(p_);
    for (int q = 0; q < 4; ++q) b2[q] = bb[q] >= 0 ? B(bb[q]) : NO_CELL;
    ed[0] = B(x.ed[UPZ]), ed[1] = B(x.ed[UMZ]), ed[2] = B(x.ed[VMM]);
    ed[3] = B(x.ed[WMP]), ed[4] = B(x.ed[WPM]), ed[5] = B(x.ed[VPP]);
    cl[0] = B(x.mp), cl[1] = B(x.pm), cl[2] = B(x.pp), cl[3] = B(x.mm);
  }
  const double* rfa = g.rf + (int64_t)qa * NCc;
  const double* rfb = g.rf + (int64_t)qb * NCc;
  const uint32_t po = B(p);
  const int y0 = chunk * FR_KC, y1 = min(y0 + FR_KC, nfl);
#pragma unroll 1
  for (int y = y0; y < y1; ++y) {
    const double* hf = hfl + (int64_t)y * 3 * g.NCf;
    const double* df = dfl + (int64_t)y * g.NCf;
    // sign * sum3 + sign * sum3d, as -sum3(..) - sum3d(..) or sum3(..) + sum3d(..)
    const double a1 = sum3i(hf, sa.a, rfa, NCc, po), a2 = sum3di(hf, sa.c, sa.d, rfa + 3 * NCc, NCc, po);
    const double sma = sga < 0.0 ? -a1 - a2 : a1 + a2;
    const double b1 = sum3i(hf, sb.a, rfb, NCc, po), b2v = sum3di(hf, sb.c, sb.d, rfb + 3 * NCc, NCc, po);
    const double smb = sgb < 0.0 ? -b1 - b2v : b1 + b2v;
    hcl[(int64_t)y * 3 * NCc + (int64_t)E * NCc + p] = flux_edge_apply(hf, df, w, ed, cl, mz, pz, b2) + sma + smb;
  }
}

__global__ __launch_bounds__(256) void ml_flux_restrict_kernel(PairGeo g, int nfl, const double* __restrict__ hfl,
                                                               const double* __restrict__ dfl,
                                                               double* __restrict__ hcl) {
  int lz, ip, jp, chunk;
  if (!parent_of(g, lz, ip, jp, chunk, 0)) return;
  if (blockIdx.z == 0)
    flux_restrict_edge<0>(g, nfl, hfl, dfl, hcl, lz, ip, jp, chunk);
  else if (blockIdx.z == 1)
    flux_restrict_edge<1>(g, nfl, hfl, dfl, hcl, lz, ip, jp, chunk);
  else
    flux_restrict_edge<2>(g, nfl, hfl, dfl, hcl, lz, ip, jp, chunk);
}

// coarse scalar trends (cal_scalar_trend, ops.f90:1325-1343) of the
// restricted fluxes over the coarse compute list (halo copies included, so
// the trends leave ghost filled for the next restriction), and the RK
// update of owned cells and pole masters: out = q + h * trend (or the trend)
__global__ void ml_coarse_scalar_kernel(int n, int nmaster, const int64_t* __restrict__ ncl, int S, int64_t NC, int L,
                                        const double* __restrict__ hinv, const double* __restrict__ hcl,
                                        double* __restrict__ dsl, const double* __restrict__ qm,
                                        const double* __restrict__ qt, double* __restrict__ om,
                                        double* __restrict__ ot, double h, int trend_only, int plus_div) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = blockIdx.y, k = y % L;
  const double* hf = hcl + (int64_t)y * 3 * NC;
  const int64_t dst = ncl[2 * t], c = ncl[2 * t + 1];
  const int64_t cW = c - 1, cS = c - S, cSW = c - S - 1;
  const double div = (hf[c] - hf[cW] + hf[NC + cSW] - hf[NC + c] + hf[2 * NC + c] - hf[2 * NC + cS]) * hinv[c];
  const double d = plus_div ? div : -div;
  dsl[(int64_t)y * NC + dst] = d;
  if (t < nmaster && om) {
    const int64_t x = (int64_t)k * NC + dst;
    if (y < L)
      om[x] = trend_only ? d : qm[x] + h * d;
    else
      ot[x] = trend_only ? d : qt[x] + h * d;
  }
}

// coarse velocity source = sum of its two child edges (u_source_cpt_restr,
// multi_level.f90:372-396), then du_grad (ops.f90:787-832) with the
// sub-sampled Bernoulli and Exner functions of the finest level (scale = its
// cells per coarse cell) and the RK update of the owned edges
__global__ void ml_coarse_velo_kernel(PairGeo g, int L, int scale, int Sfine, int64_t NCfine,
                                      const double* __restrict__ bern, const double* __restrict__ exn,
                                      const double* __restrict__ srcf, double* __restrict__ srcc,
                                      const double* __restrict__ ym, const double* __restrict__ yt,
                                      const double* __restrict__ qu, double* __restrict__ ou, double h,
                                      int trend_only) {
  int lz, ip, jp, k;
  if (!parent_of(g, lz, ip, jp, k, L)) return;
  const int64_t NC = g.NCc, NF = g.NCf;
  const double* sf = srcf + (int64_t)k * 3 * NF;
  double* sc = srcc + (int64_t)k * 3 * NC;
  const int64_t p = cellc(g, lz, ip, jp);
  const int64_t id = cellf(g, lz, 2 * ip, 2 * jp), idE = cellf(g, lz, 2 * ip + 1, 2 * jp),
                idNE = cellf(g, lz, 2 * ip + 1, 2 * jp + 1), idN = cellf(g, lz, 2 * ip, 2 * jp + 1);
  const double s0 = sf[id] + sf[idE], s1 = sf[NF + id] + sf[NF + idNE], s2 = sf[2 * NF + id] + sf[2 * NF + idN];
  sc[p] = s0;
  sc[NC + p] = s1;
  sc[2 * NC + p] = s2;
  auto fine = [&](int i, int j) {
    return (int64_t)k * NCfine + (int64_t)lz * Sfine * Sfine + (int64_t)(scale * j + 2) * Sfine + (scale * i + 2);
  };
  const double b0 = bern[fine(ip, jp)], bE = bern[fine(ip + 1, jp)], bNE = bern[fine(ip + 1, jp + 1)],
               bN = bern[fine(ip, jp + 1)];
  const double x0 = exn[fine(ip, jp)], xE = exn[fine(ip + 1, jp)], xNE = exn[fine(ip + 1, jp + 1)],
               xN = exn[fine(ip, jp + 1)];
  const int64_t cE = cellc(g, lz, ip + 1, jp), cNE = cellc(g, lz, ip + 1, jp + 1), cN = cellc(g, lz, ip, jp + 1);
  const double* m = ym + (int64_t)k * NC;
  const double* t = yt + (int64_t)k * NC;
  const double th0 = t[p] / m[p], thE = t[cE] / m[cE], thNE = t[cNE] / m[cNE], thN = t[cN] / m[cN];
  const double the[3] = {0.5 * (th0 + thE), 0.5 * (th0 + thNE), 0.5 * (th0 + thN)};
  const double* len = g.len_c;
  const double gB[3] = {(bE - b0) / len[p], (b0 - bNE) / len[NC + p], (bN - b0) / len[2 * NC + p]};
  const double gX[3] = {(xE - x0) / len[p], (x0 - xNE) / len[NC + p], (xN - x0) / len[2 * NC + p]};
  const double src[3] = {s0, s1, s2};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double du = src[e] / len[e * NC + p] - gB[e] - the[e] * gX[e];
    const int64_t x = ((int64_t)k * 3 + e) * NC + p;
    ou[x] = trend_only ? du : qu[x] + h * du;
  }
}

// cpt_or_restr_eta (barotropic_2d.f90:417-455): the coarse copy of the free
// surface takes the finer copy's value at the coincident node on [0, n]^2
__global__ void ml_eta_subsample_kernel(int nc, int Sc, int Sf, const double* __restrict__ ef,
                                        double* __restrict__ ec) {
  const int w = nc + 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)10 * w * w) return;
  const int lz = (int)(t / ((int64_t)w * w)), r = (int)(t % ((int64_t)w * w));
  const int i = r % w, j = r / w;
  ec[(int64_t)lz * Sc * Sc + (int64_t)(j + 2) * Sc + (i + 2)] =
      ef[(int64_t)lz * Sf * Sf + (int64_t)(2 * j + 2) * Sf + (2 * i + 2)];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int mlfail(wtg_ml* m, int code, const char* fmt, const char* a = "") {
  return wtg_fail(m ? m->lv[m->nl - 1] : nullptr, code, fmt, a);
}

template <class Tp>
int ml_alloc(wtg_ml* m, Tp** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, count * sizeof(Tp) > 0 ? count * sizeof(Tp) : 1) != hipSuccess)
    return wtg_fail(m->lv[m->nl - 1], WTG_E_DEVICE, "hipMalloc of %zu bytes failed", count * sizeof(Tp));
  m->allocs.push_back(q);
  *p = (Tp*)q;
  return 0;
}

template <class Tp>
int ml_upload(wtg_ml* m, Tp** p, const std::vector<Tp>& v) {
  if (int rc = ml_alloc(m, p, v.size())) return rc;
  if (hipMemcpy(*p, v.data(), v.size() * sizeof(Tp), hipMemcpyHostToDevice) != hipSuccess)
    return wtg_fail(m->lv[m->nl - 1], WTG_E_DEVICE, "upload of %zu bytes failed", v.size() * sizeof(Tp));
  return 0;
}

MgGeo mgeo(const wtg_ml* m, int l) {
  const wtg_ctx *f = m->lv[l], *c = m->lv[l - 1];
  return MgGeo{m->ovl[l], f->d_hinv, c->d_hinv, f->info.ncells, f->info.S, f->info.n, c->info.S,
               f->info.south_pole, f->info.north_pole, c->info.south_pole, c->info.north_pole, c->info.ncells};
}

PairGeo pgeo(const wtg_ml* m, int l) {
  const wtg_ctx *f = m->lv[l], *c = m->lv[l - 1];
  PairGeo g{};
  g.nc = c->info.n;
  g.Sc = c->info.S;
  g.nf = f->info.n;
  g.Sf = f->info.S;
  g.NCc = c->info.ncells;
  g.NCf = f->info.ncells;
  g.len_c = c->d_len;
  g.tri_c = c->d_tri;
  g.len_f = f->d_len;
  g.tri_f = f->d_tri;
  g.iu = m->iu[l];
  g.ovl = m->ovl[l];
  g.split = m->split[l];
  g.hinv_f = f->d_hinv;
  g.rf = m->rf[l - 1];
  g.wide = m->wide[l];
  g.WH = m->WH;
  return g;
}

#define MLCHK(m, call)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return wtg_fail((m)->lv[(m)->nl - 1], WTG_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

hipStream_t mstream(const wtg_ml* m) { return m->lv[m->nl - 1]->stream; }

// ghost fill of every level's state in one slot
int ghost_all(wtg_ml* m, const std::vector<const wtg_ctx::Slot*>& s) {
  for (int l = 0; l < m->nl; ++l)
    if (int rc = wtg_ghost_fill_layers(m->lv[l], m->L, s[l]->m, s[l]->t, s[l]->u)) return rc;
  return 0;
}

// wavelets of level l (>= 1) of a state against level l - 1
int wavelets(wtg_ml* m, int l, const wtg_ctx::Slot& fine, const wtg_ctx::Slot& coarse) {
  wtg_ctx* f = m->lv[l];
  const int L = m->L;
  const MgGeo g = mgeo(m, l);
  hipLaunchKernelGGL(ml_scalar_wavelet_kernel, dim3(nblk(f->nncl), ml_chunks(2 * L)), dim3(NT), 0, mstream(m), f->nncl, f->d_ncl,
                     g, L, fine.m, fine.t, m->wcs[l]);
  MLCHK(m, hipGetLastError());
  const PairGeo pg = pgeo(m, l);
  hipLaunchKernelGGL(ml_velo_wavelet_kernel, velo_grid(pg.nc, L), dim3(NT), 0,
                     mstream(m), pg, L,
                     coarse.u, fine.u, m->wcu[l]);
  MLCHK(m, hipGetLastError());
  hipLaunchKernelGGL(ml_velo_penta_kernel, dim3(nblk(10 * L)), dim3(NT), 0, mstream(m), pg, L, 1, coarse.u, m->wcu[l]);
  MLCHK(m, hipGetLastError());
  return wtg_ghost_fill_layers(f, L, nullptr, nullptr, m->wcu[l]);
}

// inverse transform of level l from level l - 1 and the kept wavelets
int inverse(wtg_ml* m, int l, const wtg_ctx::Slot& fine, const wtg_ctx::Slot& coarse) {
  wtg_ctx *f = m->lv[l], *c = m->lv[l - 1];
  const int L = m->L;
  const MgGeo g = mgeo(m, l);
  hipLaunchKernelGGL(ml_prolong_even_kernel, dim3(nblk(f->nncl), ml_chunks(2 * L)), dim3(NT), 0, mstream(m), f->nncl, f->d_ncl, g,
                     L, coarse.m, coarse.t, c->info.ncells, m->wcs[l], fine.m, fine.t);
  MLCHK(m, hipGetLastError());
  hipLaunchKernelGGL(ml_prolong_odd_kernel, dim3(nblk(f->nncl), ml_chunks(2 * L)), dim3(NT), 0, mstream(m), f->nncl, f->d_ncl, g,
                     L, m->wcs[l], fine.m, fine.t);
  MLCHK(m, hipGetLastError());
  const PairGeo pg = pgeo(m, l);
  const dim3 pgrid = velo_grid(pg.nc, L);
  hipLaunchKernelGGL(ml_velo_outer_kernel, pgrid, dim3(NT), 0, mstream(m), pg, L, coarse.u, m->wcu[l], fine.u);
  MLCHK(m, hipGetLastError());
  hipLaunchKernelGGL(ml_velo_penta_kernel, dim3(nblk(10 * L)), dim3(NT), 0, mstream(m), pg, L, -1, coarse.u, fine.u);
  MLCHK(m, hipGetLastError());
  if (int rc = wtg_ghost_fill_layers(f, L, nullptr, nullptr, fine.u)) return rc;
  hipLaunchKernelGGL(ml_velo_inner_kernel, pgrid, dim3(NT), 0, mstream(m), pg, L, coarse.u, m->wcu[l], fine.u);
  MLCHK(m, hipGetLastError());
  return wtg_ghost_fill_layers(f, L, nullptr, nullptr, fine.u);
}

// WT_after_step on the state in s (every level): wavelets of every level
// from the state, then the inverse transform coarse to fine
int wt_after_step(wtg_ml* m, const std::vector<const wtg_ctx::Slot*>& s) {
  if (int rc = ghost_all(m, s)) return rc;
  for (int l = 1; l < m->nl; ++l)
    if (int rc = wavelets(m, l, *s[l], *s[l - 1])) return rc;
  for (int l = 1; l < m->nl; ++l)
    if (int rc = inverse(m, l, *s[l], *s[l - 1])) return rc;
  return 0;
}

// one stage: trend of y on every level and out = q + h * trend (trend_only:
// out = trend); y ghost filled
int stage(wtg_ml* m, const std::vector<const wtg_ctx::Slot*>& y, const std::vector<const wtg_ctx::Slot*>& q,
          const std::vector<const wtg_ctx::Slot*>& out, double h, int trend_only) {
  const int nl = m->nl, top = nl - 1, L = m->L;
  wtg_ctx* ft = m->lv[top];
  const int64_t NT_ = ft->info.ncells;
  WtgStageKeep keep{};
  if (nl > 1) {
    keep.hm = m->hflux[top];
    keep.ht = m->hflux[top] + (int64_t)L * 3 * NT_;
    keep.bern = m->bern;
    keep.exn = m->exn;
    keep.src = m->src[top];
    keep.dm = m->dsc[top];
    keep.dt = m->dsc[top] + (int64_t)L * NT_;
  }
  if (int rc = wtg_launch_stage_keep(ft, *y[top], *q[top], *out[top], h, trend_only, nl > 1 ? &keep : nullptr))
    return rc;
  if (nl == 1) return 0;
  // the finest fluxes and scalar trends leave ghost filled (update_bdry of
  // horiz_flux and dq before the restriction)
  if (int rc = wtg_ghost_fill_layers(ft, 2 * L, nullptr, nullptr, m->hflux[top])) return rc;
  if (int rc = wtg_ghost_fill_layers(ft, L, keep.dm, keep.dt, nullptr)) return rc;
  for (int l = top - 1; l >= 0; --l) {
    wtg_ctx* c = m->lv[l];
    const PairGeo pg = pgeo(m, l + 1);
    const dim3 pgrid(nblk((int64_t)10 * pg.nc * pg.nc), (2 * L + FR_KC - 1) / FR_KC, 3);
    hipLaunchKernelGGL(ml_flux_restrict_kernel, pgrid, dim3(NT), 0, mstream(m), pg, 2 * L, m->hflux[l + 1], m->dsc[l + 1],
                       m->hflux[l]);
    MLCHK(m, hipGetLastError());
    if (int rc = wtg_ghost_fill_layers(c, 2 * L, nullptr, nullptr, m->hflux[l])) return rc;
    const int64_t NC = c->info.ncells;
    hipLaunchKernelGGL(ml_coarse_scalar_kernel, dim3(nblk(c->nncl), 2 * L), dim3(NT), 0, mstream(m), c->nncl,
                       c->nmaster, c->d_ncl, c->info.S, NC, L, c->d_hinv, m->hflux[l], m->dsc[l], q[l]->m, q[l]->t,
                       out[l]->m, out[l]->t, h, trend_only, 0);
    MLCHK(m, hipGetLastError());
    hipLaunchKernelGGL(ml_coarse_velo_kernel, dim3(nblk((int64_t)10 * pg.nc * pg.nc), L), dim3(NT), 0, mstream(m), pg, L,
                       1 << (top - l), ft->info.S, NT_, m->bern, m->exn, m->src[l + 1], m->src[l], y[l]->m, y[l]->t,
                       q[l]->u, out[l]->u, h, trend_only);
    MLCHK(m, hipGetLastError());
  }
  return 0;
}

int check(wtg_ml* m, int slot) {
  if (!m) return wtg_fail(nullptr, WTG_E_ARG, "null hierarchy");
  for (int l = 0; l < m->nl; ++l) {
    wtg_ctx* c = m->lv[l];
    if (slot < 0 || slot >= c->nslots) return wtg_fail(c, WTG_E_ARG, "slot %d out of range", slot);
    if (!c->topo_set) return wtg_fail(m->lv[m->nl - 1], WTG_E_STATE, "level %d: topography not set", l);
    if (c->precision != WTG_PRECISION_FP64)
      return wtg_fail(m->lv[m->nl - 1], WTG_E_STATE, "the multi-level step runs in FP64 (level %d)", l);
    if (c->dif_on)
      return wtg_fail(m->lv[m->nl - 1], WTG_E_STATE,
                      "level %d: diffusion is supported on single levels only (compute-or-restrict of its "
                      "fluxes over a hierarchy is not built)", l);
  }
  return 0;
}

std::vector<const wtg_ctx::Slot*> slots(wtg_ml* m, int which) {
  std::vector<const wtg_ctx::Slot*> v(m->nl);
  for (int l = 0; l < m->nl; ++l) v[l] = &m->lv[l]->slots[which];
  return v;
}

}  // namespace

int wtg_ml_create(const wtg_grid* const* grids, wtg_ctx* const* levels, int nlevels, wtg_ml** out) {
  if (!grids || !levels || !out) return wtg_fail(nullptr, WTG_E_ARG, "null argument");
  *out = nullptr;
  if (nlevels < 1 || nlevels > 16) return wtg_fail(nullptr, WTG_E_ARG, "nlevels %d out of range [1,16]", nlevels);
  for (int l = 0; l < nlevels; ++l) {
    if (!grids[l] || !levels[l]) return wtg_fail(nullptr, WTG_E_ARG, "level %d: null grid or context", l);
    const wtg_ctx* c = levels[l];
    if (c->info.n != grids[l]->lv.n) return wtg_fail(nullptr, WTG_E_ARG, "level %d: context and grid differ", l);
    if (l && grids[l]->lv.level != grids[l - 1]->lv.level + 1)
      return wtg_fail(nullptr, WTG_E_ARG, "levels %d and %d are not consecutive", l - 1, l);
    if (c->nmembers != 1) return wtg_fail(nullptr, WTG_E_ARG, "level %d is an ensemble context", l);
    if (c->params.nlayers != levels[0]->params.nlayers || c->nslots != levels[0]->nslots)
      return wtg_fail(nullptr, WTG_E_ARG, "level %d: layers or slots differ from level 0", l);
    if (c->device != levels[0]->device) return wtg_fail(nullptr, WTG_E_ARG, "level %d is on another device", l);
    for (int k = 0; k < l; ++k)
      if (levels[k] == c) return wtg_fail(nullptr, WTG_E_ARG, "level %d repeats a context", l);
  }
  if (nlevels > 1 && levels[nlevels - 1]->stage_version != 3)
    return wtg_fail(nullptr, WTG_E_ARG, "the finest level must use the layer-pair stage kernel");
  auto m = std::make_unique<wtg_ml>();
  m->nl = nlevels;
  m->L = levels[0]->params.nlayers;
  m->lv.assign(levels, levels + nlevels);
  m->ovl.assign(nlevels, nullptr);
  m->split.assign(nlevels, nullptr);
  m->iu.assign(nlevels, nullptr);
  m->rf.assign(nlevels, nullptr);
  m->wide.assign(nlevels, nullptr);
  m->hflux.assign(nlevels, nullptr);
  m->dsc.assign(nlevels, nullptr);
  m->src.assign(nlevels, nullptr);
  m->wcs.assign(nlevels, nullptr);
  m->wcu.assign(nlevels, nullptr);
  m->eg.assign(nlevels, nullptr);
  wtg_ml* mp = m.get();
  auto bail = [&](int rc) {
    g_create_error = mp->lv[mp->nl - 1]->err;
    wtg_ml_destroy(m.release());
    return rc;
  };
  const int L = m->L;
  for (int l = 0; l < nlevels; ++l) {
    const wtg::Level& g = grids[l]->lv;
    const size_t NC = g.ncells();
    int rc = 0;
    if (l) {
      try {
        wtg::TransferWeights w = wtg::transfer_weights(g, grids[l - 1]->lv);
        std::vector<double> a = wtg::overlap_areas(g, grids[l - 1]->lv);
        if ((rc = ml_upload(mp, &m->ovl[l], a)) || (rc = ml_upload(mp, &m->split[l], w.split)) ||
            (rc = ml_upload(mp, &m->iu[l], w.iu)) || (rc = ml_upload(mp, &m->rf[l - 1], w.rf)) ||
            (rc = ml_upload(mp, &m->wide[l], g.wide_src)))
          return bail(rc);
      } catch (const std::exception& e) {
        wtg_ml_destroy(m.release());
        return wtg_fail(nullptr, WTG_E_ARG, "%s", e.what());
      }
      m->WH = g.WH;
      if ((rc = ml_alloc(mp, &m->wcs[l], 2 * L * NC)) || (rc = ml_alloc(mp, &m->wcu[l], 3 * L * NC))) return bail(rc);
    }
    if ((rc = ml_alloc(mp, &m->hflux[l], 6 * L * NC)) || (rc = ml_alloc(mp, &m->dsc[l], 2 * L * NC)) ||
        (rc = ml_alloc(mp, &m->src[l], 3 * L * NC)) || (rc = ml_alloc(mp, &m->eg[l], NC)))
      return bail(rc);
    // fluxes outside the stage kernel's region and never-filled slots stay zero
    if (hipMemset(m->hflux[l], 0, 6 * L * NC * sizeof(double)) != hipSuccess ||
        hipMemset(m->dsc[l], 0, 2 * L * NC * sizeof(double)) != hipSuccess ||
        hipMemset(m->src[l], 0, 3 * L * NC * sizeof(double)) != hipSuccess)
      return bail(wtg_fail(mp->lv[nlevels - 1], WTG_E_DEVICE, "work buffer clear failed"));
    if (l && (hipMemset(m->wcs[l], 0, 2 * L * NC * sizeof(double)) != hipSuccess ||
              hipMemset(m->wcu[l], 0, 3 * L * NC * sizeof(double)) != hipSuccess))
      return bail(wtg_fail(mp->lv[nlevels - 1], WTG_E_DEVICE, "work buffer clear failed"));
  }
  const size_t NTop = grids[nlevels - 1]->lv.ncells();
  if (int rc = ml_alloc(mp, &m->bern, L * NTop)) return bail(rc);
  if (int rc = ml_alloc(mp, &m->exn, L * NTop)) return bail(rc);
  // every level's work is ordered on the finest level's stream
  for (int l = 0; l < nlevels; ++l) {
    m->saved_stream.push_back(levels[l]->stream);
    if (hipStreamSynchronize(levels[l]->stream) != hipSuccess)
      return bail(wtg_fail(mp->lv[nlevels - 1], WTG_E_DEVICE, "stream synchronisation failed"));
    levels[l]->stream = levels[nlevels - 1]->stream;
  }
  *out = m.release();
  return 0;
}

void wtg_ml_destroy(wtg_ml* m) {
  if (!m) return;
  if (!m->lv.empty()) hipStreamSynchronize(m->lv[m->nl - 1]->stream);
  for (size_t l = 0; l < m->saved_stream.size(); ++l) m->lv[l]->stream = m->saved_stream[l];
  for (void* p : m->allocs) hipFree(p);
  delete m;
}

int wtg_ml_forward(wtg_ml* m, int slot) {
  if (int rc = check(m, slot)) return rc;
  const auto s = slots(m, slot);
  const int L = m->L;
  for (int l = m->nl - 1; l >= 1; --l) {
    wtg_ctx *f = m->lv[l], *c = m->lv[l - 1];
    if (int rc = wtg_ghost_fill_layers(f, L, s[l]->m, s[l]->t, nullptr)) return rc;
    const MgGeo g = mgeo(m, l);
    hipLaunchKernelGGL(ml_scalar_wavelet_kernel, dim3(nblk(f->nncl), ml_chunks(2 * L)), dim3(NT), 0, mstream(m), f->nncl,
                       f->d_ncl, g, L, s[l]->m, s[l]->t, m->wcs[l]);
    MLCHK(m, hipGetLastError());
    hipLaunchKernelGGL(ml_restrict_scalar_kernel, dim3(nblk(c->nncl), ml_chunks(2 * L)), dim3(NT), 0, mstream(m), c->nncl,
                       c->d_ncl, g, L, s[l]->m, s[l]->t, m->wcs[l], s[l - 1]->m, s[l - 1]->t, c->info.ncells);
    MLCHK(m, hipGetLastError());
    if (int rc = wtg_restrict_velo(f, slot, c, slot)) return rc;
  }
  for (int l = 0; l < m->nl; ++l)
    if (int rc = wtg_ghost_fill_layers(m->lv[l], L, nullptr, nullptr, s[l]->u)) return rc;
  for (int l = 1; l < m->nl; ++l) {
    const PairGeo pg = pgeo(m, l);
    hipLaunchKernelGGL(ml_velo_wavelet_kernel, velo_grid(pg.nc, L), dim3(NT), 0,
                       mstream(m), pg, L, s[l - 1]->u, s[l]->u, m->wcu[l]);
    MLCHK(m, hipGetLastError());
    hipLaunchKernelGGL(ml_velo_penta_kernel, dim3(nblk(10 * L)), dim3(NT), 0, mstream(m), pg, L, 1, s[l - 1]->u,
                       m->wcu[l]);
    MLCHK(m, hipGetLastError());
    if (int rc = wtg_ghost_fill_layers(m->lv[l], L, nullptr, nullptr, m->wcu[l])) return rc;
  }
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  return 0;
}

int wtg_ml_inverse(wtg_ml* m, int slot) {
  if (int rc = check(m, slot)) return rc;
  const auto s = slots(m, slot);
  for (int l = 0; l < m->nl; ++l)
    if (int rc = wtg_ghost_fill_layers(m->lv[l], m->L, s[l]->m, s[l]->t, s[l]->u)) return rc;
  for (int l = 1; l < m->nl; ++l)
    if (int rc = inverse(m, l, *s[l], *s[l - 1])) return rc;
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  return 0;
}

int wtg_ml_wavelets(wtg_ml* m, int level, int to_device, double* mass, double* temp, double* velo) {
  if (!m) return wtg_fail(nullptr, WTG_E_ARG, "null hierarchy");
  if (level < 1 || level >= m->nl) return mlfail(m, WTG_E_ARG, "wavelets live on levels 1..nlevels-1");
  const int64_t NC = m->lv[level]->info.ncells, L = m->L;
  const size_t bs = (size_t)L * NC * sizeof(double);
  const hipMemcpyKind kd = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  if (mass) MLCHK(m, to_device ? hipMemcpy(m->wcs[level], mass, bs, kd) : hipMemcpy(mass, m->wcs[level], bs, kd));
  if (temp)
    MLCHK(m, to_device ? hipMemcpy(m->wcs[level] + L * NC, temp, bs, kd) : hipMemcpy(temp, m->wcs[level] + L * NC, bs, kd));
  if (velo) MLCHK(m, to_device ? hipMemcpy(m->wcu[level], velo, 3 * bs, kd) : hipMemcpy(velo, m->wcu[level], 3 * bs, kd));
  return 0;
}

int wtg_ml_trend(wtg_ml* m, int q_slot, int dq_slot) {
  if (int rc = check(m, q_slot)) return rc;
  if (int rc = check(m, dq_slot)) return rc;
  if (q_slot == dq_slot) return mlfail(m, WTG_E_ARG, "the trend needs its own slot");
  const auto q = slots(m, q_slot), d = slots(m, dq_slot);
  if (int rc = ghost_all(m, q)) return rc;
  if (int rc = stage(m, q, q, d, 0.0, 1)) return rc;
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  return 0;
}

int wtg_ml_wt_after_step(wtg_ml* m, int slot) {
  if (int rc = check(m, slot)) return rc;
  if (int rc = wt_after_step(m, slots(m, slot))) return rc;
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  return 0;
}

int wtg_ml_rk4(wtg_ml* m, int slot, double dt, int nsteps) {
  if (int rc = check(m, slot)) return rc;
  if (nsteps < 0) return mlfail(m, WTG_E_ARG, "nsteps < 0");
  const int nl = m->nl, ns = m->lv[0]->nslots;
  const auto sol = slots(m, slot), w1 = slots(m, ns), w2 = slots(m, ns + 1);
  for (int l = 0; l < nl; ++l) {
    if (int rc = wtg_copy_frozen_slots(m->lv[l], *sol[l], *w1[l])) return rc;
    if (int rc = wtg_copy_frozen_slots(m->lv[l], *sol[l], *w2[l])) return rc;
  }
  if (int rc = ghost_all(m, sol)) return rc;
  const double coef[4] = {dt / 4.0, dt / 3.0, dt / 2.0, dt};
  for (int s = 0; s < nsteps; ++s) {
    const std::vector<const wtg_ctx::Slot*>* ys[4] = {&sol, &w1, &w2, &w1};
    const std::vector<const wtg_ctx::Slot*>* os[4] = {&w1, &w2, &w1, &sol};
    for (int st = 0; st < 4; ++st) {
      if (int rc = stage(m, *ys[st], sol, *os[st], coef[st], 0)) return rc;
      if (int rc = wt_after_step(m, *os[st])) return rc;
    }
  }
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  return 0;
}

namespace {

// flux_divergence (barotropic_2d.f90:326-366) of a slot on every level into
// the field `out` of each level: the finest level's layer-integrated fluxes
// and divergence, coarser levels from the restricted flux (cpt_or_restr_flux
// with the finer divergence as dscalar), each divergence ghost filled
int flux_divergence_ml(wtg_ml* m, int slot, int out) {
  const int top = m->nl - 1;
  wtg_ctx* ft = m->lv[top];
  if (int rc = wtg_fs_flux_div_into(ft, slot, m->hflux[top], ft->fs[out])) return rc;
  if (int rc = wtg_ghost_fill_layers(ft, 1, nullptr, nullptr, m->hflux[top])) return rc;
  for (int l = top - 1; l >= 0; --l) {
    wtg_ctx* c = m->lv[l];
    const PairGeo pg = pgeo(m, l + 1);
    hipLaunchKernelGGL(ml_flux_restrict_kernel, dim3(nblk((int64_t)10 * pg.nc * pg.nc), 1, 3), dim3(NT), 0, mstream(m), pg,
                       1, m->hflux[l + 1], (const double*)m->lv[l + 1]->fs[out], m->hflux[l]);
    MLCHK(m, hipGetLastError());
    if (int rc = wtg_ghost_fill_layers(c, 1, nullptr, nullptr, m->hflux[l])) return rc;
    hipLaunchKernelGGL(ml_coarse_scalar_kernel, dim3(nblk(c->nncl), 1), dim3(NT), 0, mstream(m), c->nncl, c->nmaster,
                       c->d_ncl, c->info.S, c->info.ncells, 1, c->d_hinv, m->hflux[l], c->fs[out],
                       (const double*)nullptr, (const double*)nullptr, (double*)nullptr, (double*)nullptr, 0.0, 1, 1);
    MLCHK(m, hipGetLastError());
  }
  return 0;
}

// the free surface each level's external pressure gradient uses: the
// finest level's own, coarser levels a copy with [0, n]^2 sub-sampled from
// the next finer copy (grad_eta, barotropic_2d.f90:457-474)
int eta_for_gradient(wtg_ml* m) {
  const int top = m->nl - 1;
  m->eg[top] = nullptr;
  for (int l = top - 1; l >= 0; --l) {
    wtg_ctx *c = m->lv[l], *f = m->lv[l + 1];
    MLCHK(m, hipMemcpyAsync(m->eg[l], c->fs[FS_ETA], (size_t)c->info.ncells * sizeof(double), hipMemcpyDeviceToDevice,
                            mstream(m)));
    const double* ef = l + 1 == top ? f->fs[FS_ETA] : m->eg[l + 1];
    const int w = c->info.n + 1;
    hipLaunchKernelGGL(ml_eta_subsample_kernel, dim3(nblk((int64_t)10 * w * w)), dim3(NT), 0, mstream(m), c->info.n,
                       c->info.S, f->info.S, ef, m->eg[l]);
    MLCHK(m, hipGetLastError());
  }
  return 0;
}

const double* grad_field(wtg_ml* m, int l) { return l == m->nl - 1 ? m->lv[l]->fs[FS_ETA] : m->eg[l]; }

}  // namespace

int wtg_ml_rk_split(wtg_ml* m, wtg_mg* mg, int slot, int nstages, int nsteps, const wtg_mg_params* p, int* iters) {
  if (int rc = check(m, slot)) return rc;
  if (!mg || !p || !iters) return mlfail(m, WTG_E_ARG, "null argument");
  if (nsteps < 0) return mlfail(m, WTG_E_ARG, "nsteps < 0");
  if (nstages < 1 || nstages > 4) return mlfail(m, WTG_E_ARG, "split schemes have 1 to 4 stages");
  const int nl = m->nl, top = nl - 1, ns = m->lv[0]->nslots;
  if (!wtg_mg_same_levels(mg, m->lv.data(), nl))
    return mlfail(m, WTG_E_ARG, "the solver hierarchy must run on the same contexts");
  for (int l = 0; l < nl; ++l) {
    wtg_ctx* c = m->lv[l];
    if (!c->fs_params_set) return wtg_fail(m->lv[top], WTG_E_STATE, "level %d: free-surface parameters not set", l);
    if (!c->fs[FS_ETA]) return wtg_fail(m->lv[top], WTG_E_STATE, "level %d: free-surface fields not allocated", l);
    if (c->fs_dt != m->lv[top]->fs_dt || c->fs_theta1 != m->lv[top]->fs_theta1 ||
        c->fs_theta2 != m->lv[top]->fs_theta2)
      return wtg_fail(m->lv[top], WTG_E_ARG, "level %d: free-surface parameters differ from the finest level's", l);
  }
  const auto sol = slots(m, slot), w1 = slots(m, ns), w2 = slots(m, ns + 1);
  for (int l = 0; l < nl; ++l) {
    if (int rc = wtg_copy_frozen_slots(m->lv[l], *sol[l], *w1[l])) return rc;
    if (int rc = wtg_copy_frozen_slots(m->lv[l], *sol[l], *w2[l])) return rc;
  }
  const double dt = m->lv[top]->fs_dt, th2 = m->lv[top]->fs_theta2;
  // stage input, trend buffer and output, as the single-level split step:
  // sol -> q1 -> ... -> sol with h/n, ..., h/2, h
  double coef[4];
  const std::vector<const wtg_ctx::Slot*>*ins[4], *trs[4], *outs[4];
  for (int k = 0; k < nstages; ++k) {
    coef[k] = dt / (double)(nstages - k);
    outs[k] = k == nstages - 1 ? &sol : (k % 2 == 0 ? &w1 : &w2);
    ins[k] = k == 0 ? &sol : outs[k - 1];
    trs[k] = outs[k] != &sol ? outs[k] : (ins[k] == &w1 ? &w2 : &w1);
  }
  *iters = 0;
  for (int s = 0; s < nsteps; ++s) {
    if (int rc = ghost_all(m, sol)) return rc;
    for (int l = 0; l < nl; ++l) {
      wtg_ctx* c = m->lv[l];
      if (int rc = wtg_ghost_fill_layers(c, 1, c->fs[FS_ETA], c->fs[FS_ETA], nullptr)) return rc;
    }
    if (th2 != 1.0)
      if (int rc = flux_divergence_ml(m, slot, FS_DIVF_OLD)) return rc;
    if (int rc = eta_for_gradient(m)) return rc;
    for (int st = 0; st < nstages; ++st) {
      const auto &in = *ins[st], &tr = *trs[st], &out = *outs[st];
      for (int l = 0; l < nl; ++l)
        if (int rc = wtg_baro_correction_slot(m->lv[l], *in[l])) return rc;
      if (int rc = ghost_all(m, in)) return rc;
      if (int rc = stage(m, in, in, tr, 0.0, 1)) return rc;
      for (int l = 0; l < nl; ++l)
        if (int rc = wtg_split_axpy(m->lv[l], coef[st], grad_field(m, l), *sol[l], *tr[l], *out[l])) return rc;
      if (int rc = wt_after_step(m, out)) return rc;
    }
    // free_surface_update: eta_update, barotropic_correction, u_update,
    // WT_after_step with the free surface
    if (int rc = flux_divergence_ml(m, slot, FS_DIVF)) return rc;
    for (int l = 0; l < nl; ++l) {
      wtg_ctx* c = m->lv[l];
      if (int rc = wtg_fs_rhs_async(c)) return rc;
      MLCHK(m, hipMemcpyAsync(c->fs[FS_ETA_OLD], c->fs[FS_ETA], (size_t)c->info.ncells * sizeof(double),
                              hipMemcpyDeviceToDevice, mstream(m)));
    }
    // FMG reports the iterations and the residual of every level
    std::vector<int> its(nl, 0);
    std::vector<double> res(nl, 0.0);
    if (int rc = wtg_mg_solve(mg, WTG_SOLVE_FMG, p, its.data(), res.data())) return rc;
    for (int l = 0; l < nl; ++l) *iters += its[l];
    for (int l = 0; l < nl; ++l)
      if (int rc = wtg_baro_correction_slot(m->lv[l], *sol[l])) return rc;
    if (int rc = eta_for_gradient(m)) return rc;
    for (int l = 0; l < nl; ++l)
      if (int rc = wtg_u_update(m->lv[l], grad_field(m, l), *sol[l])) return rc;
    if (int rc = wtg_mg_eta_transform(mg)) return rc;
    if (int rc = wt_after_step(m, sol)) return rc;
  }
  MLCHK(m, hipStreamSynchronize(mstream(m)));
  return 0;
}

int wtg_ml_rk4_split(wtg_ml* m, wtg_mg* mg, int slot, int nsteps, const wtg_mg_params* p, int* iters) {
  return wtg_ml_rk_split(m, mg, slot, 4, nsteps, p, iters);
}
