// Icosahedral lozenge grid: coordinates, ghost (halo) maps and TRiSK geometry
// for one uniform level, in the lozenge-tile layout used by every kernel.
//
// Layout.  The sphere is 10 lozenges (pairs of icosahedron faces) of n x n
// nodes, n = 2^level, plus the two poles.  Each lozenge is stored as a tile of
// S x S cells, S = n + 2*HALO, covering lozenge-local (i, j) in
// [-HALO, n+HALO); cell (lz, i, j) has index lz*S*S + (j+HALO)*S + (i+HALO).
// Owned nodes are i, j in [0, n); the south pole lives in halo cell (n, 0) of
// lozenge 0 and the north pole in halo cell (0, n) of lozenge 5 (their
// "master" cells), which is where the reference keeps them (init.f90:547-569,
// pole_master).  Every node owns three edges RT (i,j)->(i+1,j), DG
// (i+1,j+1)->(i,j), UP (i,j)->(i,j+1) and two triangles LORT
// {(i,j),(i+1,j),(i+1,j+1)} and UPLT {(i,j),(i+1,j+1),(i,j+1)}
// (shared.f90:134-149, Users_manual.md:683-690); edge/triangle quantities are
// planes of the same cell indexing.
//
// Halo cells hold copies of owned values (scalars) or +-copies (edges, sign
// -1 where the neighbour lozenge orients the edge the other way), as the
// reference's pack/unpk ghost lists do (comm.f90:103-529, 785-861).  The map
// is derived here geometrically (the lattice continues across lozenge edges),
// and at the icosahedron vertices by the reference's conventions.
//
// Geometry follows init.f90:247-311 (corners, midpoint subdivision),
// init.f90:651-820 and sphere.f90:15-294 (circumcentres, lengths, hexagon
// part areas, triangle areas, Coriolis), and refine_patch.f90:134-186.
#include "grid.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <array>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace wtg {

namespace {

struct V3 {
  double x, y, z;
};

inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double nrm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// sphere.f90:91-101 project_on_sphere
inline V3 project(V3 p, double R) {
  double s = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
  return {p.x * (R / s), p.y * (R / s), p.z * (R / s)};
}
// sphere.f90:249-254 mid_pt
inline V3 mid_pt(V3 p, V3 q, double R) { return project(add(p, q), R); }

inline V3 cross(V3 u, V3 v) {
  return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}

// sphere.f90:15-21 dist
inline double dist(V3 p, V3 q, double R) {
  double a = p.y * q.z - p.z * q.y, b = p.z * q.x - p.x * q.z, c = p.x * q.y - p.y * q.x;
  return R * std::asin(std::sqrt(a * a + b * b + c * c) / (R * R));
}
// sphere.f90:186-198 distn
inline double distn(V3 p, V3 q, double R) {
  double a = p.y * q.z - p.z * q.y, b = p.z * q.x - p.x * q.z, c = p.x * q.y - p.y * q.x;
  double s = (1.0 / R) * (1.0 / R) * std::sqrt(a * a + b * b + c * c);
  if (s > 1.0) return std::asin(1.0);
  return std::asin(s);
}
// sphere.f90:163-184 triarea (L'Huilier)
inline double triarea(V3 A, V3 B, V3 C, double R) {
  double ab = distn(A, B, R), ac = distn(A, C, R), bc = distn(B, C, R);
  double s = (ab + ac + bc) / 2;
  double t = std::tan(0.5 * s) * std::tan((s - ab) / 2.0) * std::tan((s - ac) / 2.0) *
             std::tan((s - bc) / 2.0);
  if (t < 1e-64) return 0.0;
  return 4.0 * R * R * std::atan(std::sqrt(t));
}
// sphere.f90:200-213 circumcentre
inline V3 circumcentre(V3 A, V3 B, V3 C, double R) {
  V3 c = cross(sub(A, B), sub(C, B));
  if (nrm(c) < R * 1e-13) return c;
  return project(c, R);
}
// sphere.f90:80-86 sph2cart
inline V3 sph2cart(double lon, double lat, double R) {
  return {R * (std::cos(lon) * std::cos(lat)), R * (std::sin(lon) * std::cos(lat)), R * std::sin(lat)};
}

struct Key {
  int64_t a, b, c;
  bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    return (size_t)(k.a * 73856093) ^ (size_t)(k.b * 19349663) ^ (size_t)(k.c * 83492791);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------
Level build_level(int level, int min_level, int domain_level, double radius, double omega) {
  const int patch_size = 1 << (min_level - domain_level - 1);
  if (level < 1 || level > 12) throw std::runtime_error("build_level: level out of range");
  Level g;
  g.level = level;
  g.n = 1 << level;
  g.S = g.n + 2 * HALO;
  g.radius = radius;
  g.omega = omega;
  const int n = g.n, S = g.S;
  const double R = radius;
  const size_t NC = g.ncells();
  std::vector<V3> X(NC, V3{0, 0, 0});
  auto cell = [&](int lz, int i, int j) -> size_t { return g.cell(lz, i, j); };

  // --- 1. lozenge corners (init.f90:260-287) -------------------------------
  const double PI = std::acos(-1.0);
  double lat[4] = {-PI / 2.0, -std::atan(1.0 / 2.0), std::atan(1.0 / 2.0), PI / 2.0};
  double lon[10];
  for (int k = 0; k < 10; ++k) lon[k] = (PI * (double)k) / 5.0;
  auto modp = [](int a, int m) { return ((a % m) + m) % m; };

  // per-lozenge natural lattice on [0, n]^2.  Sub-domain corners follow
  // get_J0_coord (init.f90:289-300): recursive midpoints of the unit-sphere
  // lozenge corners down to domain_level, each projected once more by
  // assign_coord (init.f90:651-696); inside a sub-domain every new node is the
  // midpoint of the two coarser nodes of the edge it splits (assign_coord and
  // refine_patch2, refine_patch.f90:154-163), which is what makes the
  // coordinates bit-identical to the reference's.
  if (domain_level < 0 || min_level - 1 < domain_level || level < min_level - 1)
    throw std::runtime_error("build_level: need domain_level <= min_level-1 <= level");
  const int ND = 1 << domain_level;
  // root level min_level-1: one patch per sub-domain
  int m = 1 << (min_level - 1);
  const int N = m / ND;
  std::vector<std::vector<V3>> nat(10, std::vector<V3>((size_t)(m + 1) * (m + 1)));
  for (int ii = 1; ii <= 2; ++ii)
    for (int jj = 1; jj <= 5; ++jj) {
      int lz = 5 * ii - 5 + (jj - 1);
      // Fortran 1-based lon(m) = pi*(m-1)/5, lat(m); corners on the unit sphere
      // as sph2cart(...)/radius
      auto unit = [&](double lo, double la) {
        V3 c = sph2cart(lo, la, R);
        return V3{c.x / R, c.y / R, c.z / R};
      };
      V3 cnr[2][2];
      cnr[1][1] = unit(lon[modp(ii + 2 * jj - 2, 10)], lat[ii]);      // ne
      cnr[1][0] = unit(lon[(ii + 2 * jj - 2) - 1], lat[ii - 1]);      // se
      cnr[0][0] = unit(lon[modp(ii + 2 * jj - 4, 10)], lat[ii]);      // sw
      cnr[0][1] = unit(lon[(ii + 2 * jj - 2) - 1], lat[ii + 1]);      // nw
      std::function<V3(int, int, int)> j0 = [&](int i, int j, int l) -> V3 {
        if (l > 0) return mid_pt(j0(i / 2, j / 2, l - 1), j0((i + 1) / 2, (j + 1) / 2, l - 1), R);
        return cnr[i][j];
      };
      auto& L = nat[lz];
      auto at = [&](int i, int j) -> V3& { return L[(size_t)j * (m + 1) + i]; };
      for (int dj = 0; dj <= ND; ++dj)
        for (int di = 0; di <= ND; ++di) at(di * N, dj * N) = project(j0(di, dj, domain_level), R);
      for (int step = N; step > 1; step /= 2) {
        int h = step / 2;
        for (int j = 0; j <= m; j += step)
          for (int i = 0; i <= m; i += step) {
            if (i + step <= m) at(i + h, j) = mid_pt(at(i, j), at(i + step, j), R);
            if (j + step <= m) at(i, j + h) = mid_pt(at(i, j), at(i, j + step), R);
            if (i + step <= m && j + step <= m)
              at(i + h, j + h) = mid_pt(at(i + step, j + step), at(i, j), R);
          }
      }
    }
  // Ghost nodes carry their owner's coordinates (comm_nodes3), and each finer
  // level is made of the coarser nodes and the midpoints of coarser edges
  // computed from those ghost-filled coordinates (refine_patch2).
  auto make_consistent = [&](int mm) {
    const double tol = 1e-9 * R;
    const double qq = R * 1.1 / (double)mm / 8.0;
    auto kk = [&](V3 p) { return Key{(int64_t)std::floor(p.x / qq), (int64_t)std::floor(p.y / qq), (int64_t)std::floor(p.z / qq)}; };
    std::unordered_map<Key, std::vector<std::array<int, 3>>, KeyHash> own;
    for (int lz = 0; lz < 10; ++lz)
      for (int j = 0; j < mm; ++j)
        for (int i = 0; i < mm; ++i) own[kk(nat[lz][(size_t)j * (mm + 1) + i])].push_back({lz, i, j});
    V3 spole = nat[0][(size_t)0 * (mm + 1) + mm], npole = nat[5][(size_t)mm * (mm + 1) + 0];
    for (int lz = 0; lz < 10; ++lz)
      for (int j = 0; j <= mm; ++j)
        for (int i = 0; i <= mm; ++i) {
          if (i < mm && j < mm) continue;
          V3& p = nat[lz][(size_t)j * (mm + 1) + i];
          if (nrm(sub(p, spole)) < tol) { p = spole; continue; }
          if (nrm(sub(p, npole)) < tol) { p = npole; continue; }
          Key k = kk(p);
          bool done = false;
          for (int a = -1; a <= 1 && !done; ++a)
            for (int b = -1; b <= 1 && !done; ++b)
              for (int c = -1; c <= 1 && !done; ++c) {
                auto it = own.find(Key{k.a + a, k.b + b, k.c + c});
                if (it == own.end()) continue;
                for (auto& o : it->second) {
                  V3 q2 = nat[o[0]][(size_t)o[2] * (mm + 1) + o[1]];
                  if (nrm(sub(q2, p)) < tol) { p = q2; done = true; break; }
                }
              }
          if (!done) throw std::runtime_error("build_level: boundary node without owner");
        }
  };
  make_consistent(m);
  while (m < n) {
    int mc = 2 * m;
    std::vector<std::vector<V3>> fine(10, std::vector<V3>((size_t)(mc + 1) * (mc + 1)));
    for (int lz = 0; lz < 10; ++lz) {
      auto P = [&](int i, int j) { return nat[lz][(size_t)j * (m + 1) + i]; };
      auto& F = fine[lz];
      auto at = [&](int i, int j) -> V3& { return F[(size_t)j * (mc + 1) + i]; };
      for (int j = 0; j <= m; ++j)
        for (int i = 0; i <= m; ++i) {
          at(2 * i, 2 * j) = P(i, j);
          if (i < m) at(2 * i + 1, 2 * j) = mid_pt(P(i, j), P(i + 1, j), R);
          if (j < m) at(2 * i, 2 * j + 1) = mid_pt(P(i, j), P(i, j + 1), R);
          if (i < m && j < m) at(2 * i + 1, 2 * j + 1) = mid_pt(P(i + 1, j + 1), P(i, j), R);
        }
    }
    nat.swap(fine);
    m = mc;
    make_consistent(m);
  }

  // --- 2. owned nodes, poles and a coordinate hash ---------------------------
  const double h_est = R * 1.1 / (double)n;  // < lattice spacing (~1.1 R/n)
  const double q = h_est / 8.0;
  std::unordered_map<Key, int64_t, KeyHash> owner;
  auto key = [&](V3 p) {
    return Key{(int64_t)std::floor(p.x / q), (int64_t)std::floor(p.y / q), (int64_t)std::floor(p.z / q)};
  };
  auto insert = [&](V3 p, int64_t c) {
    Key k = key(p);
    owner[k] = c;
  };
  auto lookup = [&](V3 p) -> int64_t {
    Key k = key(p);
    int64_t best = -1;
    double bd = 1e300;
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          auto it = owner.find(Key{k.a + a, k.b + b, k.c + c});
          if (it == owner.end()) continue;
          double d = nrm(sub(X[it->second], p));
          if (d < bd) { bd = d; best = it->second; }
        }
    if (best < 0 || bd > 0.25 * h_est) return -1;
    return best;
  };
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        size_t c = cell(lz, i, j);
        X[c] = nat[lz][(size_t)j * (n + 1) + i];
        insert(X[c], (int64_t)c);
      }
  g.south_pole = cell(0, n, 0);
  g.north_pole = cell(5, 0, n);
  X[g.south_pole] = nat[0][(size_t)0 * (n + 1) + n];
  X[g.north_pole] = nat[5][(size_t)n * (n + 1) + 0];
  insert(X[g.south_pole], (int64_t)g.south_pole);
  insert(X[g.north_pole], (int64_t)g.north_pole);

  // --- 3. halo cells: source cell of every halo node --------------------------
  // Every lattice point on a lozenge boundary also belongs to the natural
  // lattice of the neighbouring lozenge(s).  A halo cell (lz, i, j) is reached
  // from the clamped anchor (ic, jc) in [0, n]^2 by an offset d; in a
  // neighbour's frame the same anchor sits at (i', j') and the hexagonal
  // neighbour directions are rotated by r sixths of a turn (found from an
  // along-edge neighbour), so the halo cell is (i', j') + rot_r(d) there.
  static const int cyc[6][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 0}, {-1, -1}, {0, -1}};
  auto dir_index = [&](int a, int b) {
    for (int k = 0; k < 6; ++k)
      if (cyc[k][0] == a && cyc[k][1] == b) return k;
    return -1;
  };
  // hex-lattice rotation: decompose d on the two basis vectors (1,0), (0,1)
  auto rot = [&](int r, int a, int b, int& ra, int& rb) {
    const int* e1 = cyc[(0 + r) % 6];
    const int* e2 = cyc[(2 + r) % 6];
    ra = a * e1[0] + b * e2[0];
    rb = a * e1[1] + b * e2[1];
  };
  std::unordered_multimap<Key, int64_t, KeyHash> natpts;  // -> lz*(n+1)^2 + j*(n+1) + i
  const int64_t NN = (int64_t)(n + 1) * (n + 1);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i) natpts.emplace(key(nat[lz][(size_t)j * (n + 1) + i]), lz * NN + (int64_t)j * (n + 1) + i);
  auto same = [&](V3 a, V3 b) { return nrm(sub(a, b)) < 1e-9 * R; };
  auto nat_of = [&](int lz, int i, int j) { return nat[lz][(size_t)j * (n + 1) + i]; };
  // all (lz', i', j') with lz' != lz at the same point as nat[lz](i, j)
  auto others = [&](int lz, int i, int j) {
    std::vector<std::array<int, 3>> out;
    V3 p = nat_of(lz, i, j);
    Key k = key(p);
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          auto rng = natpts.equal_range(Key{k.a + a, k.b + b, k.c + c});
          for (auto it = rng.first; it != rng.second; ++it) {
            int64_t v = it->second;
            int l2 = (int)(v / NN), rem = (int)(v % NN);
            if (l2 == lz) continue;
            int j2 = rem / (n + 1), i2 = rem % (n + 1);
            if (same(nat_of(l2, i2, j2), p)) out.push_back({l2, i2, j2});
          }
        }
    return out;
  };
  auto owned_cell_at = [&](V3 p) -> int64_t { return lookup(p); };

  // node sources on a virtual frame one cell wider than the tile, so that
  // both end nodes of every tile edge are known
  const int EH = HALO + 1, ES = n + 2 * EH;
  auto ecell = [&](int lz, int i, int j) { return (size_t)lz * ES * ES + (size_t)(j + EH) * ES + (size_t)(i + EH); };
  // owned cell at lozenge-frame point (lz, i, j) off the tile's owned square
  // (-1 where the lattice cannot be continued across the boundary)
  auto resolve = [&](int lz, int i, int j) -> int64_t {
    int ic = std::min(std::max(i, 0), n), jc = std::min(std::max(j, 0), n);
    int di = i - ic, dj = j - jc;
    if (di == 0 && dj == 0) return owned_cell_at(nat_of(lz, ic, jc));  // on the natural boundary row/column
    // anchor strictly inside the side where possible (avoid vertices)
    int64_t found = -1;
    for (auto& o : others(lz, ic, jc)) {
      int l2 = o[0], i2 = o[1], j2 = o[2];
      // along-edge direction: a lattice neighbour of the anchor that is
      // also on the shared boundary
      for (int k = 0; k < 6 && found < 0; ++k) {
        int ai = ic + cyc[k][0], aj = jc + cyc[k][1];
        if (ai < 0 || ai > n || aj < 0 || aj > n) continue;
        V3 pa = nat_of(lz, ai, aj);
        for (int kk = 0; kk < 6; ++kk) {
          int bi = i2 + cyc[kk][0], bj = j2 + cyc[kk][1];
          if (bi < 0 || bi > n || bj < 0 || bj > n) continue;
          if (!same(nat_of(l2, bi, bj), pa)) continue;
          int r = ((kk - k) % 6 + 6) % 6;
          int ra, rb;
          rot(r, di, dj, ra, rb);
          int ti = i2 + ra, tj = j2 + rb;
          if (ti < 0 || ti > n || tj < 0 || tj > n) break;
          // the target must lie on the far side of the shared edge
          int64_t oc = owned_cell_at(nat_of(l2, ti, tj));
          if (oc >= 0) found = oc;
          break;
        }
      }
      if (found >= 0) break;
    }
    return found;
  };
  std::vector<int64_t> esrc((size_t)10 * ES * ES, -1);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -EH; j < n + EH; ++j)
      for (int i = -EH; i < n + EH; ++i) {
        size_t c = ecell(lz, i, j);
        if (i >= 0 && i < n && j >= 0 && j < n) {
          esrc[c] = (int64_t)cell(lz, i, j);
          continue;
        }
        if ((lz == 0 && i == n && j == 0) || (lz == 5 && i == 0 && j == n)) {
          esrc[c] = (int64_t)cell(lz, i, j);
          continue;
        }
        esrc[c] = resolve(lz, i, j);
      }
  // a wider frame of plain lattice continuation, for the few stencils that
  // reach more than HALO cells beyond a lozenge side (flux restriction)
  g.WH = HALO + 3;
  g.WS = n + 2 * g.WH;
  g.wide_src.assign((size_t)10 * g.WS * g.WS, -1);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -g.WH; j < n + g.WH; ++j)
      for (int i = -g.WH; i < n + g.WH; ++i) {
        const size_t c = g.wcell(lz, i, j);
        if (i >= 0 && i < n && j >= 0 && j < n) g.wide_src[c] = (int64_t)cell(lz, i, j);
        else if (i >= -EH && j >= -EH && i < n + EH && j < n + EH) g.wide_src[c] = esrc[ecell(lz, i, j)];
        else g.wide_src[c] = resolve(lz, i, j);
      }
  (void)dir_index;
  // --- 4. pole neighbourhoods ------------------------------------------------
  // Five lozenges meet at each pole with their acute corners, so the lattice
  // cannot be continued there.  The reference fills these ghost cells so that
  // the pole's missing sixth neighbour duplicates its neighbour across the
  // corner (S == E at the south pole, W == N at the north pole); the pattern is
  // invariant under the 72-degree rotation lz -> lz+1 (mod 5).
  for (int k = 0; k < 5; ++k) {
    int o = (k + 3) % 5;  // lozenge across the pole corner
    esrc[ecell(k, n, -1)] = (int64_t)cell(o, n - 1, 0);
    esrc[ecell(k, n + 1, 0)] = (int64_t)cell(o, n - 1, 0);
    esrc[ecell(k, n, -2)] = (int64_t)cell(o, n - 2, 0);
    esrc[ecell(k, n - 1, -2)] = (int64_t)cell(o, n - 1, 1);
    esrc[ecell(k, n + 1, -1)] = (int64_t)cell((k + 2) % 5, n - 1, 1);
    int kn = 5 + k, on = 5 + (k + 3) % 5;
    esrc[ecell(kn, -1, n)] = (int64_t)cell(on, 0, n - 1);
    esrc[ecell(kn, 0, n + 1)] = (int64_t)cell(on, 0, n - 1);
    esrc[ecell(kn, -2, n)] = (int64_t)cell(on, 0, n - 2);
    esrc[ecell(kn, -2, n - 1)] = (int64_t)cell(on, 1, n - 1);
    esrc[ecell(kn, -1, n + 1)] = (int64_t)cell(5 + (k + 2) % 5, 1, n - 1);
  }
  // At the other vertex corners of every lozenge the sixth neighbour is
  // missing too: (-1,-1) beyond the SW pentagon and (n+1,n+1) beyond the NE
  // pentagon hold no lattice node.  The reference leaves their values unused
  // (every stencil weight there is zero); a neighbour's value keeps them finite.
  for (int lz = 0; lz < 10; ++lz) {
    esrc[ecell(lz, -1, -1)] = esrc[ecell(lz, 0, -1)];
    esrc[ecell(lz, n + 1, n + 1)] = lz < 5 ? esrc[ecell(lz, n, n + 1)] : esrc[ecell(lz, n + 1, n)];
  }
  g.node_src.assign(NC, -1);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -HALO; j < n + HALO; ++j)
      for (int i = -HALO; i < n + HALO; ++i) g.node_src[cell(lz, i, j)] = esrc[ecell(lz, i, j)];
  const std::vector<int64_t> geo_src = esrc;  // coordinates of every cell, frozen or not

  // --- 5. edge ghost map -------------------------------------------------------
  // A halo edge joins two halo/owned nodes; its source is the owned edge with
  // the same end nodes, and the sign says whether the owner orients it the
  // same way (RT: tail (i,j) head (i+1,j); DG: tail (i+1,j+1) head (i,j);
  // UP: tail (i,j) head (i,j+1)).
  static const int tail_d[3][2] = {{0, 0}, {1, 1}, {0, 0}};
  static const int head_d[3][2] = {{1, 0}, {0, 0}, {0, 1}};
  auto owned_ij = [&](int64_t c, int& lz, int& i, int& j) {
    lz = (int)(c / ((int64_t)S * S));
    int64_t r = c % ((int64_t)S * S);
    j = (int)(r / S) - HALO;
    i = (int)(r % S) - HALO;
  };
  g.edge_src.assign(3 * NC, -1);
  g.edge_sign.assign(3 * NC, 0);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -HALO; j < n + HALO; ++j)
      for (int i = -HALO; i < n + HALO; ++i) {
        size_t c = cell(lz, i, j);
        for (int e = 0; e < 3; ++e) {
          size_t slot = 3 * c + e;
          if (i >= 0 && i < n && j >= 0 && j < n) {
            g.edge_src[slot] = (int64_t)slot;
            g.edge_sign[slot] = 1;
            continue;
          }
          int ti = i + tail_d[e][0], tj = j + tail_d[e][1];
          int hi = i + head_d[e][0], hj = j + head_d[e][1];
          int64_t a = esrc[ecell(lz, ti, tj)], b = esrc[ecell(lz, hi, hj)];
          if (a < 0 || b < 0 || a == b) continue;
          // candidate owners: either end node, in its own frame
          for (int pass = 0; pass < 2 && g.edge_src[slot] < 0; ++pass) {
            int64_t own = pass == 0 ? a : b;
            int l2, i2, j2;
            owned_ij(own, l2, i2, j2);
            if (!(i2 >= 0 && i2 < n && j2 >= 0 && j2 < n)) continue;  // pole masters own no edges
            for (int e2 = 0; e2 < 3; ++e2) {
              int oi = i2 + (e2 == 0 ? 1 : e2 == 1 ? 1 : 0), oj = j2 + (e2 == 0 ? 0 : 1);
              int64_t other = esrc[ecell(l2, oi, oj)];
              if (other < 0) continue;
              int64_t otail = (e2 == 1) ? other : own, ohead = (e2 == 1) ? own : other;
              if (otail == a && ohead == b) {
                g.edge_src[slot] = 3 * (int64_t)cell(l2, i2, j2) + e2;
                g.edge_sign[slot] = 1;
                break;
              }
              if (otail == b && ohead == a) {
                g.edge_src[slot] = 3 * (int64_t)cell(l2, i2, j2) + e2;
                g.edge_sign[slot] = -1;
                break;
              }
            }
          }
        }
      }

  // Corner edges by the reference's conventions (comm.f90:395-428, 436-529):
  // the RT edge of the artificial SW cell is the pentagon's edge
  // (-1,0)->(0,-1); the RT edge of (n, n+1) at the NE pentagon is
  // (n,n+1)->(n+1,n); beside the south pole the UP edge of (n+1, 0) copies
  // the DG edge of cell (n - patch_size, patch_size - 1) of lozenge lz+2,
  // which is what the reference's pole ghost lists deliver.
  auto set_edge_between = [&](int lz, int i, int j, int e, int64_t a, int64_t b) {
    size_t slot = 3 * cell(lz, i, j) + e;
    g.edge_src[slot] = -1;
    g.edge_sign[slot] = 0;
    for (int pass = 0; pass < 2; ++pass) {
      int64_t own = pass == 0 ? a : b;
      int l2, i2, j2;
      owned_ij(own, l2, i2, j2);
      if (!(i2 >= 0 && i2 < n && j2 >= 0 && j2 < n)) continue;
      for (int e2 = 0; e2 < 3; ++e2) {
        int oi = i2 + (e2 == 2 ? 0 : 1), oj = j2 + (e2 == 0 ? 0 : 1);
        int64_t other = esrc[ecell(l2, oi, oj)];
        int64_t otail = (e2 == 1) ? other : own, ohead = (e2 == 1) ? own : other;
        if ((otail == a && ohead == b) || (otail == b && ohead == a)) {
          g.edge_src[slot] = 3 * (int64_t)cell(l2, i2, j2) + e2;
          g.edge_sign[slot] = (otail == a) ? 1 : -1;
          return;
        }
      }
    }
  };
  for (int lz = 0; lz < 10; ++lz) {
    set_edge_between(lz, -1, -1, 0, esrc[ecell(lz, -1, 0)], esrc[ecell(lz, 0, -1)]);
    set_edge_between(lz, n, n + 1, 0, esrc[ecell(lz, n, n + 1)], esrc[ecell(lz, n + 1, n)]);
  }
  if (patch_size > 0 && patch_size < n)
    for (int k = 0; k < 5; ++k) {
      size_t slot = 3 * cell(k, n + 1, 0) + 2;
      g.edge_src[slot] = 3 * (int64_t)cell((k + 2) % 5, n - patch_size, patch_size - 1) + 1;
      g.edge_sign[slot] = 1;
    }

  // Halo cells and edges the reference's ghost lists never fill (found by
  // poisoning every ghost value before update_bdry).  They keep whatever
  // value they were initialised with for the whole run: the ghost fill skips
  // them and a time step leaves them unchanged (their trend is zero), so they
  // are "frozen" here too (source -1).  Some are read with non-zero weight
  // beside the icosahedron vertices, so callers must initialise them exactly
  // as the reference's initial state does.
  {
    const int RT = 0, DG = 1, UP = 2, ALL = 3;
    struct Z { int i, j, e; };
    const Z south[] = {{n, n + 1, DG}, {n, n + 1, UP}, {n, -1, ALL}, {n + 1, n + 1, UP}, {n + 1, -1, ALL},
                       {n + 1, 0, DG}, {n + 1, 0, RT}, {n + 1, 1, DG}, {n + 1, 1, RT}, {-1, n, ALL},
                       {-1, n + 1, ALL}, {-1, -1, DG}, {-1, -1, UP}};
    const Z north[] = {{n, -1, ALL}, {n + 1, n, ALL}, {n + 1, n + 1, RT}, {n + 1, -1, ALL}, {-1, n, ALL},
                       {-1, n + 1, ALL}, {-1, -1, DG}, {0, n + 1, DG}, {0, n + 1, UP}, {1, n + 1, DG},
                       {1, n + 1, UP}};
    const int south_nodes[][2] = {{-1, n + 1}, {-1, -1}};
    const int north_nodes[][2] = {{n + 1, -1}, {-1, -1}};
    for (int lz = 0; lz < 10; ++lz) {
      auto drop = [&](int i, int j, int e) {
        for (int k = 0; k < 3; ++k) {
          if (e != ALL && e != k) continue;
          size_t slot = 3 * cell(lz, i, j) + k;
          g.edge_src[slot] = -1;
          g.edge_sign[slot] = 0;
        }
      };
      if (lz < 5) {
        for (auto& t : south) drop(t.i, t.j, t.e);
        for (auto& t : south_nodes) esrc[ecell(lz, t[0], t[1])] = -1;
      } else {
        for (auto& t : north) drop(t.i, t.j, t.e);
        for (auto& t : north_nodes) esrc[ecell(lz, t[0], t[1])] = -1;
      }
    }
    for (int lz = 0; lz < 10; ++lz)
      for (int j = -HALO; j < n + HALO; ++j)
        for (int i = -HALO; i < n + HALO; ++i) g.node_src[cell(lz, i, j)] = esrc[ecell(lz, i, j)];
  }

  // --- 6. geometry on the virtual frame, in each lozenge's own frame ---------
  // Coordinates of halo cells are the owners' (comm_nodes3 after refinement,
  // refine_patch.f90:154-175), except the two artificial corner cells, which
  // the reference places at the circumcentre of the coarser level's pentagon
  // triangle (ccentre_penta + midpt, init.f90:313-384, 820-870).
  const size_t ENC = (size_t)10 * ES * ES;
  std::vector<V3> XE(ENC, V3{0, 0, 0});
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -EH; j < n + EH; ++j)
      for (int i = -EH; i < n + EH; ++i) {
        int64_t s0 = esrc[ecell(lz, i, j)];
        if (s0 < 0) s0 = geo_src[ecell(lz, i, j)];
        if (s0 >= 0) XE[ecell(lz, i, j)] = X[s0];
      }
  for (int lz = 0; lz < 10; ++lz) {
    auto P = [&](int i, int j) { return XE[ecell(lz, i, j)]; };
    XE[ecell(lz, -1, -1)] = circumcentre(P(0, 0), P(0, -2), P(-2, 0), R);
    XE[ecell(lz, n + 1, n + 1)] = circumcentre(P(n, n), P(n, n + 2), P(n + 2, n), R);
  }
  std::vector<V3> CL(ENC, V3{0, 0, 0}), CU(ENC, V3{0, 0, 0});
  std::vector<V3> MP(3 * ENC, V3{0, 0, 0});
  std::vector<double> LEN(3 * ENC, 0.0), PED(3 * ENC, 0.0), AR(6 * ENC, 0.0), HI(ENC, 0.0);
  std::vector<double> TA(2 * ENC, 0.0), CO(2 * ENC, 0.0);
  for (int lz = 0; lz < 10; ++lz) {
    auto e = [&](int i, int j) { return ecell(lz, i, j); };
    auto P = [&](int i, int j) { return XE[e(i, j)]; };
    // circumcentres (init.f90:707-723), cells [-3, n+1]
    for (int j = -EH; j <= n + 1; ++j)
      for (int i = -EH; i <= n + 1; ++i) {
        CL[e(i, j)] = circumcentre(P(i, j), P(i + 1, j + 1), P(i + 1, j), R);
        CU[e(i, j)] = circumcentre(P(i, j), P(i, j + 1), P(i + 1, j + 1), R);
      }
    // pentagon corners (ccentre_penta, init.f90:313-384)
    CL[e(-1, -1)] = circumcentre(P(0, 0), P(0, -1), P(-1, 0), R);
    CU[e(-1, -1)] = CL[e(-1, -1)];
    {
      V3 pt3 = mid_pt(P(0, -2), P(-2, 0), R);
      CL[e(-1, -2)] = circumcentre(P(0, -1), P(0, -2), pt3, R);
      CU[e(-2, -1)] = circumcentre(pt3, P(-2, 0), P(-1, 0), R);
    }
    CL[e(n - 1, -1)] = circumcentre(P(n - 1, -1), P(n, 0), P(n + 1, 0), R);
    CU[e(n, -1)] = CL[e(n - 1, -1)];
    CU[e(-1, n - 1)] = circumcentre(P(-1, n - 1), P(0, n + 1), P(0, n), R);
    CL[e(-1, n)] = CU[e(-1, n - 1)];
    CL[e(n, n)] = circumcentre(P(n, n), P(n, n + 1), P(n + 1, n), R);
    CU[e(n, n)] = CL[e(n, n)];
    // edge midpoints (init.f90:820-870), cells [-1, n+1]
    for (int j = -1; j <= n + 1; ++j)
      for (int i = -1; i <= n + 1; ++i) {
        size_t c = e(i, j);
        if (j >= n + 1) { MP[3 * c + 0] = mid_pt(P(i, j), P(i + 1, j), R); continue; }
        if (i >= n + 1) { MP[3 * c + 2] = mid_pt(P(i, j), P(i, j + 1), R); continue; }
        if (j == -1 && i == -1) { MP[3 * c + 1] = CL[c]; continue; }
        if (j == -1 && i == n) { MP[3 * c + 2] = CU[c]; continue; }
        if (j == n && i == -1) { MP[3 * c + 0] = CL[c]; continue; }
        MP[3 * c + 0] = mid_pt(P(i, j), P(i + 1, j), R);
        MP[3 * c + 1] = mid_pt(P(i + 1, j + 1), P(i, j), R);
        MP[3 * c + 2] = mid_pt(P(i, j), P(i, j + 1), R);
        if (j == n && i == n) MP[3 * c + 1] = CL[c];
      }
    // primal and dual edge lengths (init.f90:386-426), cells [-1, n+1]
    for (int j = -1; j <= n + 1; ++j)
      for (int i = -1; i <= n + 1; ++i) {
        size_t c = e(i, j);
        if (j >= n + 1) { LEN[3 * c + 0] = dist(P(i, j), P(i + 1, j), R); continue; }
        if (i >= n + 1) { LEN[3 * c + 2] = dist(P(i, j), P(i, j + 1), R); continue; }
        LEN[3 * c + 0] = dist(P(i, j), P(i + 1, j), R);
        LEN[3 * c + 1] = dist(P(i + 1, j + 1), P(i, j), R);
        LEN[3 * c + 2] = dist(P(i, j), P(i, j + 1), R);
        PED[3 * c + 0] = dist(CU[e(i, j - 1)], CL[c], R);
        PED[3 * c + 1] = dist(CU[c], CL[c], R);
        PED[3 * c + 2] = dist(CU[c], CL[e(i - 1, j)], R);
        if (i == -1 && j == -1) {
          PED[3 * c + 1] = 0.0;
          LEN[3 * c + 0] = dist(P(i + 1, j), P(i, j + 1), R);
        }
        if (i == n && j == n) LEN[3 * c + 1] = dist(P(i + 1, j), P(i, j + 1), R);
      }
    // hexagon part areas (init.f90:713-761, sphere.f90:280-294), cells [-1, n+1]
    for (int j = -1; j <= n + 1; ++j)
      for (int i = -1; i <= n + 1; ++i) {
        size_t c = e(i, j);
        V3 cor[6] = {CL[c], CU[c], CL[e(i - 1, j)], CU[e(i - 1, j - 1)], CL[e(i - 1, j - 1)], CU[e(i, j - 1)]};
        V3 mid[6] = {MP[3 * c + 0], MP[3 * c + 1], MP[3 * c + 2], MP[3 * e(i - 1, j) + 0],
                     MP[3 * e(i - 1, j - 1) + 1], MP[3 * e(i, j - 1) + 2]};
        double part[6], sum = 0.0;
        for (int k = 0; k < 6; ++k) {
          part[k] = triarea(P(i, j), cor[k], mid[k], R) + triarea(P(i, j), cor[k], mid[(k + 1) % 6], R);
        }
        for (int k = 0; k < 6; ++k) sum += part[k];
        int lo = 0, hi = 5;
        if (j >= n + 1) { lo = 3; hi = 5; }
        else if (i >= n + 1) { lo = 2; hi = 4; }
        for (int k = lo; k <= hi; ++k) AR[6 * c + k] = part[k];
        if (lo == 0) HI[c] = (sum != 0.0) ? 1.0 / sum : 1.0;
      }
    // triangle areas and Coriolis (init.f90:725-818), cells [-1, n+1]
    for (int j = -1; j <= n + 1; ++j)
      for (int i = -1; i <= n + 1; ++i) {
        size_t c = e(i, j);
        size_t cE = e(i + 1, j), cNE = e(i + 1, j + 1), cN = e(i, j + 1);
        double tl = AR[6 * c + 0] + AR[6 * cE + 2] + AR[6 * cNE + 4];
        double tu = AR[6 * c + 1] + AR[6 * cNE + 3] + AR[6 * cN + 5];
        TA[2 * c + 0] = tl;
        TA[2 * c + 1] = tu;
        CO[2 * c + 0] = CL[c].z / R * 2.0 * omega * tl;
        CO[2 * c + 1] = CU[c].z / R * 2.0 * omega * tu;
      }
  }
  // hex_inv of halo cells reached by the ghost lists is the owner's
  // (set_areas, refine_patch.f90:436-450); the part areas stay local.
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -1; j <= n + 1; ++j)
      for (int i = -1; i <= n + 1; ++i) {
        if (i >= 0 && i < n && j >= 0 && j < n) continue;
        size_t c = ecell(lz, i, j);
        int64_t s0 = esrc[c];
        if (s0 < 0 || s0 == (int64_t)cell(lz, i, j)) continue;
        int l2, i2, j2;
        owned_ij(s0, l2, i2, j2);
        HI[c] = HI[ecell(l2, i2, j2)];
      }

  g.EH = EH;
  g.ES = ES;
  g.xe.resize(3 * ENC);
  g.cle.resize(3 * ENC);
  g.cue.resize(3 * ENC);
  for (size_t x = 0; x < ENC; ++x) {
    g.xe[3 * x] = XE[x].x; g.xe[3 * x + 1] = XE[x].y; g.xe[3 * x + 2] = XE[x].z;
    g.cle[3 * x] = CL[x].x; g.cle[3 * x + 1] = CL[x].y; g.cle[3 * x + 2] = CL[x].z;
    g.cue[3 * x] = CU[x].x; g.cue[3 * x + 1] = CU[x].y; g.cue[3 * x + 2] = CU[x].z;
  }
  g.mpe.resize(9 * ENC);
  for (size_t x = 0; x < 3 * ENC; ++x) {
    g.mpe[3 * x] = MP[x].x; g.mpe[3 * x + 1] = MP[x].y; g.mpe[3 * x + 2] = MP[x].z;
  }

  // --- 7. copy the tile window out of the virtual frame ------------------------
  g.X.assign(3 * NC, 0.0);
  g.len.assign(3 * NC, 0.0);
  g.pedlen.assign(3 * NC, 0.0);
  g.area.assign(6 * NC, 0.0);
  g.hex_inv.assign(NC, 0.0);
  g.triarea.assign(2 * NC, 0.0);
  g.coriolis.assign(2 * NC, 0.0);
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -HALO; j < n + HALO; ++j)
      for (int i = -HALO; i < n + HALO; ++i) {
        size_t c = cell(lz, i, j), x = ecell(lz, i, j);
        g.X[c] = XE[x].x;
        g.X[NC + c] = XE[x].y;
        g.X[2 * NC + c] = XE[x].z;
        for (int k = 0; k < 3; ++k) {
          g.len[k * NC + c] = LEN[3 * x + k];
          g.pedlen[k * NC + c] = PED[3 * x + k];
        }
        for (int k = 0; k < 6; ++k) g.area[k * NC + c] = AR[6 * x + k];
        g.hex_inv[c] = HI[x];
        for (int k = 0; k < 2; ++k) {
          g.triarea[k * NC + c] = TA[2 * x + k];
          g.coriolis[k * NC + c] = CO[2 * x + k];
        }
      }
  // --- 7. reduction weights ---------------------------------------------------
  // The reference's norms and dot products (l2, dp in lin_solve.f90:433-518)
  // visit, per domain, the domain's (N+1) x (N+1) nodes including its E and N
  // boundary rows (apply_onescale_to_patch with (st,en) = (0,1),
  // domain.f90:358-383), so nodes on domain boundaries count more than once.
  // red_weight[c] is that visit count folded onto the owner of each copy.
  {
    const int N = n >> domain_level, nsd = 1 << domain_level;
    g.red_weight.assign(NC, 0.0);
    for (int lz = 0; lz < 10; ++lz)
      for (int sj = 0; sj < nsd; ++sj)
        for (int si = 0; si < nsd; ++si)
          for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
              const int I = si * N + i, J = sj * N + j;
              size_t c = cell(lz, I, J);
              if ((I >= n || J >= n) && c != g.south_pole && c != g.north_pole) {
                const int64_t o = g.node_src[c];
                if (o < 0) throw std::runtime_error("reduction sweep reaches a halo cell without owner");
                c = (size_t)o;
              }
              g.red_weight[c] += 1.0;
            }
  }
  (void)S;
  return g;
}


// ---------------------------------------------------------------------------
// Overlap areas between two levels
// ---------------------------------------------------------------------------
namespace {

inline double inner(V3 u, V3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }

// sphere.f90:102-140 arc_inters: intersection of the great-circle arcs
// (a1, a2) and (b1, b2); the point is b2 when a2 and b2 coincide
bool arc_inters(V3 a1, V3 a2, V3 b1, V3 b2, V3& pt, double R) {
  pt = b2;
  const double eps = R * 1e-13;
  if (nrm(sub(b2, a2)) < eps) return true;
  V3 n1 = cross(a1, a2);
  double inpr = inner(n1, b1) * inner(n1, b2);
  if (inpr > 0.0) return false;
  V3 n2 = cross(b1, b2);
  inpr = inner(n2, a1) * inner(n2, a2);
  if (inpr > 0.0) return false;
  pt = project(cross(n1, n2), R);
  V3 neg{-pt.x, -pt.y, -pt.z};
  if (nrm(sub(a1, neg)) < nrm(sub(a1, pt))) pt = neg;
  return true;
}

struct Tri {
  int di, dj, t;  // triangle of node (i+di, j+dj): t = 0 LORT, 1 UPLT
};
// shared.f90:308-327 (no_adj_tri, adj_tri, bfly_tri), hex_s_offs = (2, 0, 4)
constexpr Tri NO_ADJ[10] = {{0, 0, 0},   {0, -1, 1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, 0},
                            {0, 0, 1},   {0, 0, 0},  {0, -1, 1},  {-1, -1, 0}, {-1, -1, 1}};
constexpr Tri ADJ[3][2] = {{{0, -1, 1}, {0, 0, 0}}, {{0, 0, 1}, {0, 0, 0}}, {{0, 0, 1}, {-1, 0, 0}}};
constexpr Tri BFLY[3][4] = {{{-1, -1, 0}, {0, -1, 0}, {1, 0, 1}, {0, 0, 1}},
                            {{0, 1, 0}, {-1, 0, 0}, {0, -1, 1}, {1, 0, 1}},
                            {{0, 0, 0}, {0, 1, 0}, {-1, 0, 1}, {-1, -1, 1}}};
constexpr int HEX_S_OFFS[3] = {2, 0, 4};

}  // namespace

// get_overl_areas (wavelet.f90:1408-1475) for the child of coarse node
// (ip, jp) of lozenge lz on edge e: the fine node (2ip + [e<2], 2jp + [e>0])
// gets area[0..7] (a(1..4), then the split parts area(5..8)) and typ[0..1]
// (INSIDE 0, OUTER1 1, OUTER2 2 for the sides k = 0, 1).  Returns false when
// the stencil leaves the extended frames.
bool overl_node(const Level& fine, const Level& coarse, int lz, int ip, int jp, int e, double area[8], int typ[2]) {
  const double R = fine.radius;
  auto V = [](const std::vector<double>& v, size_t x) { return V3{v[3 * x], v[3 * x + 1], v[3 * x + 2]}; };
  const int nf = fine.n, nc = coarse.n;
  auto fcc = [&](int i, int j, const Tri& t) {
    size_t x = fine.ecell(lz, i + t.di, j + t.dj);
    return t.t == 0 ? V(fine.cle, x) : V(fine.cue, x);
  };
  auto ccc = [&](int i, int j, const Tri& t) {
    size_t x = coarse.ecell(lz, i + t.di, j + t.dj);
    return t.t == 0 ? V(coarse.cle, x) : V(coarse.cue, x);
  };
  const int ic = 2 * ip + (e == 0 || e == 1 ? 1 : 0), jc = 2 * jp + (e == 1 || e == 2 ? 1 : 0);
  if (ic - 1 < -fine.EH || jc - 1 < -fine.EH || ic > nf + 1 || jc > nf + 1) return false;
  if (ip - 1 < -coarse.EH || jp - 1 < -coarse.EH || ip + 1 > nc + 1 || jp + 1 > nc + 1) return false;
  V3 hex[6];
  const int so = HEX_S_OFFS[2 - e];
  for (int k = 0; k < 6; ++k) hex[k] = fcc(ic, jc, NO_ADJ[k + so]);
  // coarse circumcentres around the parent edge: tri(row, col)
  V3 tri[3][2];
  tri[0][0] = ccc(ip, jp, BFLY[e][3]);
  tri[1][0] = ccc(ip, jp, ADJ[e][1]);
  tri[2][0] = ccc(ip, jp, BFLY[e][2]);
  tri[0][1] = ccc(ip, jp, BFLY[e][1]);
  tri[1][1] = ccc(ip, jp, ADJ[e][0]);
  tri[2][1] = ccc(ip, jp, BFLY[e][0]);
  const V3 pt = V(fine.xe, fine.ecell(lz, ic, jc));
  for (int k = 0; k < 8; ++k) area[k] = 0.0;
  typ[0] = typ[1] = 0;
  area[0] = triarea(hex[5], hex[0], pt, R);
  area[1] = triarea(hex[2], hex[3], pt, R);
  for (int i = 1; i <= 2; ++i) {
    const int c = i - 1;                     // tri column
    const V3 h1 = hex[3 * i - 3], h2 = hex[3 * i - 2], h3 = hex[3 * i - 1];  // hex(3i-2), hex(3i-1), hex(3i)
    V3 p0, p1;
    bool d0 = arc_inters(tri[0][c], tri[1][c], h1, h2, p0, R);
    bool d1 = arc_inters(tri[2][c], tri[1][c], h3, h2, p1, R);
    double& ai = area[i - 1];
    double& aj = area[2 - i];   // area(-i+3)
    if (d0 && d1) {
      area[i + 3] = triarea(p0, tri[1][c], h2, R);
      area[i + 5] = triarea(tri[1][c], h2, p1, R);
      area[i + 1] = area[i + 3] + area[i + 5];
      ai = ai + triarea(h1, p0, pt, R) + triarea(p0, pt, tri[1][c], R);
      aj = aj + triarea(p1, h3, pt, R) + triarea(tri[1][c], pt, p1, R);
      typ[2 - i] = 0;
    } else if (!d0 && !d1) {
      area[i + 1] = 0.0;
      d0 = arc_inters(tri[1][0], tri[1][1], h1, h2, p0, R);
      d1 = arc_inters(tri[1][1], tri[1][0], h2, h3, p1, R);
      if (!d0 && d1) {
        ai = ai + triarea(h1, h2, pt, R) + triarea(h2, p1, pt, R);
        aj = aj + triarea(p1, h3, pt, R);
        typ[2 - i] = 2;
      } else if (d0 && !d1) {
        ai = ai + triarea(h1, p0, pt, R);
        aj = aj + triarea(h2, h3, pt, R) + triarea(p0, h2, pt, R);
        typ[2 - i] = 1;
      }
    }
  }
  return true;
}

std::vector<double> overlap_areas(const Level& fine, const Level& coarse) {
  if (fine.n != 2 * coarse.n) throw std::runtime_error("overlap_areas: levels are not consecutive");
  const size_t NC = fine.ncells();
  std::vector<double> a(4 * NC, 0.0);
  const int nf = fine.n, nc = coarse.n;
  for (int lz = 0; lz < 10; ++lz)
    // parents whose children fall in the fine tile window with a complete stencil
    for (int jp = -1; jp <= nc; ++jp)
      for (int ip = -1; ip <= nc; ++ip)
        for (int e = 0; e < 3; ++e) {
          const int ic = 2 * ip + (e == 0 || e == 1 ? 1 : 0), jc = 2 * jp + (e == 1 || e == 2 ? 1 : 0);
          if (ic < -HALO || jc < -HALO || ic >= nf + HALO || jc >= nf + HALO) continue;
          double area[8];
          int typ[2];
          if (!overl_node(fine, coarse, lz, ip, jp, e, area, typ)) continue;
          const size_t cf = fine.cell(lz, ic, jc);
          for (int k = 0; k < 4; ++k) a[k * NC + cf] = area[k];
        }
  // Halo nodes hold their owner's areas, as the reference's node
  // communication leaves them (get_areas / set_areas, refine_patch.f90:
  // 436-467): across a rotated lozenge join the midpoint changes type (RT,
  // DG, UP) and its two edge end points trade places, so a(1) <-> a(2) and
  // a(3) <-> a(4).  Nodes that no communication reaches keep zero areas, and
  // so do the nodes (n, -1) and (-1, n) beside the acute lozenge corners
  // (area_post_comm, refine_patch.f90:469-486).
  auto kind = [](int i, int j) { return (i & 1) | ((j & 1) << 1); };
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -HALO; j < nf + HALO; ++j)
      for (int i = -HALO; i < nf + HALO; ++i) {
        if (i >= 0 && i < nf && j >= 0 && j < nf) continue;
        const size_t c = fine.cell(lz, i, j);
        if (c == fine.south_pole || c == fine.north_pole) continue;
        const int64_t src = fine.node_src[c];
        const bool corner = (i == nf && j == -1) || (i == -1 && j == nf);
        if (src < 0 || corner || (i & 1) + (j & 1) == 0) {
          for (int k = 0; k < 4; ++k) a[k * NC + c] = 0.0;
          continue;
        }
        const int r = (int)(src % ((int64_t)fine.S * fine.S));
        const int si = r % fine.S - HALO, sj = r / fine.S - HALO;
        static const int same[4] = {0, 1, 2, 3}, swapped[4] = {1, 0, 3, 2};
        const int* perm = kind(i, j) == kind(si, sj) ? same : swapped;
        for (int k = 0; k < 4; ++k) a[k * NC + c] = a[perm[k] * NC + (size_t)src];
      }
  return a;
}


// ---------------------------------------------------------------------------
// Transfer weights of the velocity and flux transforms
// ---------------------------------------------------------------------------
namespace {

// 6 x 6 dense solve by LU with partial pivoting in the order of LAPACK's
// DGETF2 (pivot = first largest |a|, column scaled by the pivot's
// reciprocal, rank-1 update) and DGETRS (row swaps, unit lower then upper
// triangular solve, column-oriented as the reference BLAS DTRSM).  G is
// column-major (G[c][r]); b is overwritten by the solution.
void dgesv6(double G[6][6], double b[6]) {
  int ipiv[6];
  for (int j = 0; j < 6; ++j) {
    int jp = j;
    double amax = std::fabs(G[j][j]);
    for (int i = j + 1; i < 6; ++i)
      if (std::fabs(G[j][i]) > amax) { amax = std::fabs(G[j][i]); jp = i; }
    ipiv[j] = jp;
    if (G[j][jp] != 0.0) {
      if (jp != j)
        for (int c = 0; c < 6; ++c) std::swap(G[c][j], G[c][jp]);
      const double sfmin = 2.2250738585072014e-308;
      if (std::fabs(G[j][j]) >= sfmin) {
        const double r = 1.0 / G[j][j];
        for (int i = j + 1; i < 6; ++i) G[j][i] = r * G[j][i];
      } else {
        for (int i = j + 1; i < 6; ++i) G[j][i] = G[j][i] / G[j][j];
      }
    }
    for (int c = j + 1; c < 6; ++c) {
      if (G[c][j] != 0.0) {
        const double t = -1.0 * G[c][j];
        for (int i = j + 1; i < 6; ++i) G[c][i] = G[c][i] + G[j][i] * t;
      }
    }
  }
  for (int i = 0; i < 6; ++i)
    if (ipiv[i] != i) std::swap(b[i], b[ipiv[i]]);
  for (int k = 0; k < 6; ++k)
    if (b[k] != 0.0)
      for (int i = k + 1; i < 6; ++i) b[i] = b[i] - b[k] * G[k][i];
  for (int k = 5; k >= 0; --k)
    if (b[k] != 0.0) {
      b[k] = b[k] / G[k][k];
      for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * G[k][i];
    }
}

inline V3 scal(double a, V3 v) { return {a * v.x, a * v.y, a * v.z}; }
// sphere.f90:256-268 normalize_Coord (zero below radius * 1e-13)
inline V3 normalize(V3 v, double R) {
  const double n = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  if (n >= R * 1e-13) return {v.x / n, v.y / n, v.z / n};
  return {0, 0, 0};
}

// wavelet.f90:1761-1775 coords_to_rowd: direction (u, v) of dirvec in the
// local frame (x, y) and the midpoint's local coordinates
inline void row_d(V3 midpt, V3 dir, V3 x, V3 y, double out[6]) {
  double q0 = inner(dir, x), q1 = inner(dir, y);
  const double n = std::sqrt(q0 * q0 + q1 * q1);
  const double u = q0 / n, v = q1 / n;
  const double a = inner(midpt, x), b = inner(midpt, y);
  out[0] = u; out[1] = u * a; out[2] = u * b; out[3] = v; out[4] = v * a; out[5] = v * b;
}

// shared.f90:282-327 tables (0-based edge index e = RT 0, DG 1, UP 2)
constexpr int HEXS[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {-1, 0, 0}, {-1, -1, 1},
                             {0, -1, 2}, {0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {-1, 0, 0}};
constexpr int NGHB[10][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 0}};
constexpr int ENDPT[3][2][2] = {{{0, 0}, {1, 0}}, {{1, 1}, {0, 0}}, {{0, 0}, {0, 1}}};   // [e][k][i|j]
constexpr int OPPNO[3][2][2] = {{{0, -1}, {1, 1}}, {{0, 1}, {1, 0}}, {{1, 1}, {-1, 0}}};

// flux interpolation points (shared.f90:85-115)
enum { UZM = 0, UPZ, UMZ, UZP, VMM, VPM, VMP, VPP, WMM, WPM, WMP, WPP, WMMM, WPPP, VMPP, VPMM };

struct Frame {
  const Level& g;
  int lz;
  V3 node(int i, int j) const {
    const size_t x = g.ecell(lz, i, j);
    return {g.xe[3 * x], g.xe[3 * x + 1], g.xe[3 * x + 2]};
  }
  V3 cc(int i, int j, int t) const {
    const size_t x = g.ecell(lz, i, j);
    const std::vector<double>& v = t == 0 ? g.cle : g.cue;
    return {v[3 * x], v[3 * x + 1], v[3 * x + 2]};
  }
  V3 mid(int i, int j, int e) const {
    const size_t x = 3 * g.ecell(lz, i, j) + e;
    return {g.mpe[3 * x], g.mpe[3 * x + 1], g.mpe[3 * x + 2]};
  }
  bool in(int i, int j) const { return i >= -g.EH && j >= -g.EH && i < g.n + g.EH && j < g.n + g.EH; }
};

// outer_velo_weights (wavelet.f90:1604-1705) for coarse node (i0, j0), edge
// e0: the nine interpolation weights minus Iu_Base_Wgt
void outer_velo_weights(const Frame& C, int i0, int j0, int e0, double out[9]) {
  static const double base[9] = {16.0 / 16, -1.0 / 16, 1.0 / 16, 1.0 / 16, -1.0 / 16,
                                 -1.0 / 16, -1.0 / 16, 1.0 / 16, 1.0 / 16};
  const double R = C.g.radius;
  const int n = C.g.n;
  // coordinates with the pentagon substitutions of get_coord (wavelet.f90:1670-1703)
  auto coord = [&](int i, int j) -> V3 {
    if (i == -1 && j == -1) {
      if (e0 == 0) return C.node(-1, 0);
      if (e0 == 2) return C.node(0, -1);
      return C.node(i, j);
    }
    if (i == -1 && j == n) return C.node(0, n + 1);
    if (i == n && j == -1) return C.node(n + 1, 0);
    if (i == n + 1 && j == n + 1) return C.node(n + 1, n);
    return C.node(i, j);
  };
  const V3 e1 = C.node(i0 + ENDPT[e0][0][0], j0 + ENDPT[e0][0][1]);
  const V3 e2 = C.node(i0 + ENDPT[e0][1][0], j0 + ENDPT[e0][1][1]);
  const V3 mp = C.mid(i0, j0, e0);
  const V3 x = normalize(sub(e2, e1), R);
  const V3 y = normalize(cross(x, mp), R);
  auto row = [&](int o1i, int o1j, const int* hs, double r[6]) {
    const int i = i0 + o1i + hs[0], j = j0 + o1j + hs[1], e = hs[2];
    const V3 p1 = coord(i + ENDPT[e][0][0], j + ENDPT[e][0][1]);
    const V3 p2 = coord(i + ENDPT[e][1][0], j + ENDPT[e][1][1]);
    row_d(mid_pt(p1, p2, R), sub(p2, p1), x, y, r);
  };
  const int so = HEX_S_OFFS[e0];
  double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 1; k <= 2; ++k) {
    const int* ek = ENDPT[e0][k - 1];
    const int* ek2 = ENDPT[e0][2 - k];   // end_pt(:, -k+3)
    const int* ok = OPPNO[e0][k - 1];
    double G[6][6], r1[6], r2[6];
    const int zero[3] = {0, 0, e0};
    row(0, 0, zero, G[0]);
    row(ek2[0], ek2[1], HEXS[so + 2 + 3 * k - 3], G[1]);
    row(ek[0], ek[1], HEXS[so + 3 - (3 * k - 3)], G[2]);
    row(ok[0], ok[1], HEXS[so + 1 + 3 * k - 3], r1);
    row(ek[0], ek[1], HEXS[so + 2 + 3 * k - 3], r2);
    for (int m = 0; m < 6; ++m) G[3][m] = r1[m] - r2[m];
    row(ek2[0], ek2[1], HEXS[so + 3 - (3 * k - 3)], r1);
    row(ok[0], ok[1], HEXS[so + 4 - (3 * k - 3)], r2);
    for (int m = 0; m < 6; ++m) G[4][m] = r1[m] - r2[m];
    row(ek[0], ek[1], HEXS[so + 5 - (3 * k - 3)], r1);
    row(ek2[0], ek2[1], HEXS[so + 0 + 3 * k - 3], r2);
    for (int m = 0; m < 6; ++m) G[5][m] = r1[m] - r2[m];
    double b[6];
    row_d(mid_pt(mp, e2, R), sub(e2, e1), x, y, b);
    dgesv6(G, b);
    w[0] = w[0] + b[0];
    w[2 * k - 1] = w[2 * k - 1] + b[1];
    w[2 * k] = w[2 * k] + b[2];
    w[2 * k + 3] = b[3];
    w[2 * k + 4] = b[4];
    w[-2 * k + 5] = w[-2 * k + 5] + b[5];
    w[-2 * k + 6] = w[-2 * k + 6] - b[5];
  }
  for (int m = 0; m < 9; ++m) out[m] = 0.5 * w[m] - base[m];
}

// basic_F_restr_wgt (wavelet.f90:1176-1360) for coarse node (ip, jp) and
// edge e (RT or UP), child (i0, j0) = (2ip, 2jp): the four weight triplets
// in the order of id_enc (1..4); typ = the overlap types of the edge's child
void flux_restr_weights(const Frame& C, const Frame& F, int ip, int jp, int e, const int typ[2], double out[4][3]) {
  const double R = F.g.radius;
  const int i0 = 2 * ip, j0 = 2 * jp;
  int i, j;
  if (e == 2) { i = i0; j = j0 + 1; }
  else { i = i0 + 1; j = j0; }
  const int so = HEX_S_OFFS[e];
  int ije[16][3];
  auto set = [&](int id, int bi, int bj, int col) {
    ije[id][0] = bi + HEXS[col][0]; ije[id][1] = bj + HEXS[col][1]; ije[id][2] = HEXS[col][2];
  };
  set(UMZ, i, j, so + 1); set(UPZ, i, j, so + 4); set(WMP, i, j, so + 0);
  set(VPP, i, j, so + 5); set(WPM, i, j, so + 3); set(VMM, i, j, so + 2);
  const int mpi = i + NGHB[so + 0][0], mpj = j + NGHB[so + 0][1];
  const int ppi = i + NGHB[so + 5][0], ppj = j + NGHB[so + 5][1];
  const int pmi = i + NGHB[so + 3][0], pmj = j + NGHB[so + 3][1];
  const int mmi = i + NGHB[so + 2][0], mmj = j + NGHB[so + 2][1];
  set(VMP, mpi, mpj, so + 2); set(VMPP, mpi, mpj, so + 5); set(UZP, mpi, mpj, so + 4);
  set(WPPP, ppi, ppj, so + 0); set(WPP, ppi, ppj, so + 3);
  set(VPM, pmi, pmj, so + 5); set(VPMM, pmi, pmj, so + 2); set(UZM, pmi, pmj, so + 1);
  set(WMMM, mmi, mmj, so + 3); set(WMM, mmi, mmj, so + 0);

  // coords_to_row of basic_F_restr_wgt: edge (i, j, e) of the fine level
  auto row = [&](const int* q, V3 x, V3 y, double r[6]) {
    const int ee = q[2];
    const V3 p1 = F.node(q[0] + ENDPT[ee][0][0], q[1] + ENDPT[ee][0][1]);
    const V3 p2 = F.node(q[0] + ENDPT[ee][1][0], q[1] + ENDPT[ee][1][1]);
    const Tri& t1 = ADJ[ee][0];
    const Tri& t2 = ADJ[ee][1];
    const double ped = dist(F.cc(q[0] + t1.di, q[1] + t1.dj, t1.t), F.cc(q[0] + t2.di, q[1] + t2.dj, t2.t), R);
    row_d(mid_pt(p1, p2, R), sub(p2, p1), x, y, r);
    for (int m = 0; m < 6; ++m) r[m] = r[m] * ped;
  };
  auto solve = [&](int k1, const int* lcsd, V3 endpt_o, const int st[9], double w[6]) {
    const Tri& ta = ADJ[lcsd[2]][-k1 + 1];   // adj_tri(:, -k1+2, e)
    const V3 c1 = F.cc(lcsd[0] + ta.di, lcsd[1] + ta.dj, ta.t);
    const V3 m1 = F.mid(lcsd[0], lcsd[1], lcsd[2]);
    const V3 x = normalize(sub(m1, c1), R);
    const V3 y = normalize(cross(x, c1), R);
    double G[6][6], r1[6], r2[6];
    row(ije[st[0]], x, y, G[0]);
    row(ije[st[1]], x, y, G[1]);
    row(ije[st[2]], x, y, G[2]);
    for (int c = 0; c < 3; ++c) {
      row(ije[st[3 + 2 * c]], x, y, r1);
      row(ije[st[4 + 2 * c]], x, y, r2);
      for (int m = 0; m < 6; ++m) G[3 + c][m] = r1[m] - r2[m];
    }
    // coords_to_row_perp of (c1, endpt_o)
    const V3 mm = mid_pt(c1, endpt_o, R);
    row_d(mm, cross(sub(endpt_o, c1), mm), x, y, w);
    const double d = dist(c1, endpt_o, R);
    for (int m = 0; m < 6; ++m) w[m] = w[m] * d;
    dgesv6(G, w);
  };
  auto permute = [&](double w[6]) {
    double t[6];
    for (int m = 0; m < 6; ++m) t[m] = w[m];
    if (e == 0) { w[0] = t[1]; w[1] = t[2]; w[2] = t[0]; w[3] = t[4]; w[4] = t[5]; w[5] = t[3]; }
    else { w[0] = t[2]; w[1] = t[0]; w[2] = t[1]; w[3] = t[5]; w[4] = t[3]; w[5] = t[4]; }
  };
  const double eps = R * 1e-13;
  // side k = 1
  {
    const Tri& tp = ADJ[e][1];
    const V3 cp = C.cc(ip + tp.di, jp + tp.dj, tp.t);
    const int* q = ije[UZP];
    const Tri& tq = ADJ[q[2]][0];
    if (dist(cp, F.cc(q[0] + tq.di, q[1] + tq.dj, tq.t), R) < eps) {
      for (int m = 0; m < 3; ++m) out[0][m] = out[1][m] = 0.0;
    } else {
      const int* lcsd = typ[1] == 1 ? ije[VPP] : typ[1] == 2 ? ije[WMP] : ije[UZP];
      static const int st[9] = {VPP, WMP, UZP, UPZ, WPP, VMP, UMZ, WPPP, VMPP};
      double w[6];
      solve(1, lcsd, cp, st, w);
      permute(w);
      for (int m = 0; m < 3; ++m) { out[0][m] = w[m]; out[1][m] = w[3 + m]; }
    }
  }
  // side k = 0
  {
    const Tri& tp = ADJ[e][0];
    const V3 cp = C.cc(ip + tp.di, jp + tp.dj, tp.t);
    const int* q = ije[UZM];
    const Tri& tq = ADJ[q[2]][1];
    if (dist(cp, F.cc(q[0] + tq.di, q[1] + tq.dj, tq.t), R) < eps) {
      for (int m = 0; m < 3; ++m) out[2][m] = out[3][m] = 0.0;
    } else {
      const int* lcsd = typ[0] == 1 ? ije[VMM] : typ[0] == 2 ? ije[WPM] : ije[UZM];
      static const int st[9] = {VMM, WPM, UZM, UMZ, WMM, VPM, UPZ, WMMM, VPMM};
      double w[6];
      solve(0, lcsd, cp, st, w);
      permute(w);
      for (int m = 0; m < 3; ++m) { out[2][m] = w[m]; out[3][m] = w[3 + m]; }
    }
  }
}

}  // namespace

TransferWeights transfer_weights(const Level& fine, const Level& coarse) {
  if (fine.n != 2 * coarse.n) throw std::runtime_error("transfer_weights: levels are not consecutive");
  TransferWeights w;
  const size_t NF = fine.ncells(), NCc = coarse.ncells();
  w.nc = (int)NCc;
  w.split.assign(2 * NF, 0.0);
  w.iu.assign(9 * NF, 0.0);
  w.rf.assign(24 * NCc, 0.0);
  const int nc = coarse.n, nf = fine.n;
  for (int lz = 0; lz < 10; ++lz) {
    const Frame C{coarse, lz}, F{fine, lz};
    for (int jp = -1; jp <= nc; ++jp)
      for (int ip = -1; ip <= nc; ++ip)
        for (int e = 0; e < 3; ++e) {
          const int ic = 2 * ip + (e == 0 || e == 1 ? 1 : 0), jc = 2 * jp + (e == 1 || e == 2 ? 1 : 0);
          if (ic < -HALO || jc < -HALO || ic >= nf + HALO || jc >= nf + HALO) continue;
          double area[8];
          int typ[2];
          if (!overl_node(fine, coarse, lz, ip, jp, e, area, typ)) continue;
          const size_t cf = fine.cell(lz, ic, jc);
          w.split[cf] = area[4];
          w.split[NF + cf] = area[5];
          const bool owned = ip >= 0 && jp >= 0 && ip < nc && jp < nc;
          if (!owned) continue;
          // I_u_wgt of the outer child edge whose second end point is the
          // child (set_WT_wgts, wavelet.f90:1585-1602)
          const int i2 = 2 * ip + ENDPT[e][1][0], j2 = 2 * jp + ENDPT[e][1][1];
          double iu[9];
          outer_velo_weights(C, ip, jp, e, iu);
          const size_t c2 = fine.cell(lz, i2, j2);
          for (int m = 0; m < 9; ++m) w.iu[m * NF + c2] = iu[m];
          if (e == 1) continue;
          double rf[4][3];
          flux_restr_weights(C, F, ip, jp, e, typ, rf);
          const size_t cp = coarse.cell(lz, ip, jp);
          const int s0 = e == 0 ? 0 : 2;   // RT: sets 0 (k=0), 1 (k=1); UP: 2, 3
          // id_enc (3, 4) belong to side k = 0, (1, 2) to side k = 1
          for (int m = 0; m < 3; ++m) {
            w.rf[((s0 + 0) * 6 + m) * NCc + cp] = rf[2][m];
            w.rf[((s0 + 0) * 6 + 3 + m) * NCc + cp] = rf[3][m];
            w.rf[((s0 + 1) * 6 + m) * NCc + cp] = rf[0][m];
            w.rf[((s0 + 1) * 6 + 3 + m) * NCc + cp] = rf[1][m];
          }
        }
  }
  // halo split parts: the owner's, unswapped (set_areas, refine_patch.f90:436-450)
  for (int lz = 0; lz < 10; ++lz)
    for (int j = -HALO; j < nf + HALO; ++j)
      for (int i = -HALO; i < nf + HALO; ++i) {
        if (i >= 0 && i < nf && j >= 0 && j < nf) continue;
        const size_t c = fine.cell(lz, i, j);
        if (c == fine.south_pole || c == fine.north_pole) continue;
        const int64_t src = fine.node_src[c];
        if (src < 0 || (i & 1) + (j & 1) == 0) {
          w.split[c] = w.split[NF + c] = 0.0;
          continue;
        }
        w.split[c] = w.split[(size_t)src];
        w.split[NF + c] = w.split[NF + (size_t)src];
      }
  return w;
}

}  // namespace wtg
