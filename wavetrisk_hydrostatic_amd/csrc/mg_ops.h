// Internal: device helpers of the scalar wavelet transform between two
// uniform levels, shared by the multigrid solvers and the multi-level step.
// Child cells of coarse cell (i, j) are, on the next finer level and in the
// same lozenge frame, the coincident node (2i, 2j) and the edge midpoints
// (2i+1, 2j) RT, (2i+1, 2j+1) DG, (2i, 2j+1) UP.  Overlap areas a(1..4) live
// at the fine midpoints (four planes); halo copies carry their owner's areas.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct MgGeo {
  const double* ovl;     // [4][NCf]
  const double* hinv_f;  // fine 1/area
  const double* hinv_c;  // coarse 1/area
  int64_t NCf;
  int Sf, nf, Sc;
  int64_t sp_f, np_f, sp_c, np_c;  // pole masters
  int64_t NCc;                     // coarse cells (member stride of coarse fields)
};

__device__ __forceinline__ void mg_ij(int64_t c, int S, int& lz, int& i, int& j) {
  const int64_t T = (int64_t)S * S;
  lz = (int)(c / T);
  const int r = (int)(c - (int64_t)lz * T);
  i = r % S - 2;
  j = r / S - 2;
}

// Interp_node (wavelet.f90:1487-1498) at fine midpoint x from its four
// coincident neighbours: the edge's end points and the two opposite nodes
template <class F>
__device__ __forceinline__ double mg_interp(const MgGeo& g, const F* __restrict__ s, int64_t x, int i, int j) {
  const int64_t S = g.Sf, NC = g.NCf;
  int64_t id1, id2, id3, id4;
  if ((i & 1) && !(j & 1)) {  // RT
    id1 = x - 1; id2 = x + 1; id3 = x + 1 + 2 * S; id4 = x - 1 - 2 * S;
  } else if ((i & 1) && (j & 1)) {  // DG
    id1 = x + 1 + S; id2 = x - 1 - S; id3 = x + 1 - S; id4 = x - 1 + S;
  } else {  // UP
    id1 = x - S; id2 = x + S; id3 = x - 2 - S; id4 = x + 2 + S;
  }
  return (g.ovl[x] * s[id1] + g.ovl[NC + x] * s[id2] + g.ovl[2 * NC + x] * s[id3] + g.ovl[3 * NC + x] * s[id4]) *
         g.hinv_f[x];
}

// lifting around fine coincident cell c (Restrict_scalar wavelet.f90:542-567,
// twelve terms; Prolong_scalar 586-611 omits id2NE)
template <class F>
__device__ __forceinline__ double mg_lift(const MgGeo& g, const F* __restrict__ wc, int64_t c, bool with_2ne) {
  const int64_t S = g.Sf, NC = g.NCf;
  const double *a1 = g.ovl, *a2 = g.ovl + NC, *a3 = g.ovl + 2 * NC, *a4 = g.ovl + 3 * NC;
  const int64_t E = c + 1, NE = c + 1 + S, N2E = c + 2 + S, N2NE = c + 1 + 2 * S, N = c + S, W = c - 1,
                NW = c - 1 + S, S2W = c - 2 - S, SW = c - 1 - S, So = c - S, SW2 = c - 1 - 2 * S, SE = c + 1 - S;
  if (with_2ne)
    return wc[E] * a1[E] + wc[NE] * a2[NE] + wc[N2E] * a3[N2E] + wc[N2NE] * a4[N2NE] + wc[N] * a1[N] +
           wc[W] * a2[W] + wc[NW] * a3[NW] + wc[S2W] * a4[S2W] + wc[SW] * a1[SW] + wc[So] * a2[So] +
           wc[SW2] * a3[SW2] + wc[SE] * a4[SE];
  return wc[E] * a1[E] + wc[NE] * a2[NE] + wc[N2E] * a3[N2E] + wc[N] * a1[N] + wc[W] * a2[W] + wc[NW] * a3[NW] +
         wc[S2W] * a4[S2W] + wc[SW] * a1[SW] + wc[So] * a2[So] + wc[SW2] * a3[SW2] + wc[SE] * a4[SE];
}

// lifting at a pole: its five edge midpoints (area a2, the pole being their
// second end point) and the five DG midpoints of the second ring (a3 south,
// a4 north); the prolongation drops the second-ring term of lozenge 1 (south)
// / 7 (north), the pole's id2NE
template <class F>
__device__ __forceinline__ double mg_pole_lift(const MgGeo& g, const F* __restrict__ wc, bool south,
                                               bool prolong) {
  const int64_t S = g.Sf, NC = g.NCf, T = S * S;
  const int n = g.nf;
  double sum = 0.0;
  if (south) {
    for (int k = 0; k < 5; ++k) {
      const int64_t m = k * T + 2 * S + (n - 1 + 2), d = k * T + 3 * S + (n - 1 + 2);
      sum += wc[m] * g.ovl[NC + m];
      if (!(prolong && k == 1)) sum += wc[d] * g.ovl[2 * NC + d];
    }
  } else {
    for (int k = 5; k < 10; ++k) {
      const int64_t m = k * T + (int64_t)(n - 1 + 2) * S + 2, d = k * T + (int64_t)(n - 1 + 2) * S + 3;
      sum += wc[m] * g.ovl[NC + m];
      if (!(prolong && k == 7)) sum += wc[d] * g.ovl[3 * NC + d];
    }
  }
  return sum;
}

