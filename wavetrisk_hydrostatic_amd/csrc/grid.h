// Lozenge-tile grid of one uniform icosahedral level (see grid.cpp).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace wtg {

constexpr int HALO = 2;      // halo cells on every side of a lozenge tile
constexpr int NLOZ = 10;     // lozenges of the icosahedron

struct Level {
  int level = 0, n = 0, S = 0;
  double radius = 0, omega = 0;
  size_t south_pole = 0, north_pole = 0;  // master cells of the poles

  size_t ncells() const { return (size_t)NLOZ * S * S; }
  size_t cell(int lz, int i, int j) const {
    return (size_t)lz * S * S + (size_t)(j + HALO) * S + (size_t)(i + HALO);
  }

  // node_src[c]: owned cell whose value halo cell c copies (c itself for owned
  // cells and pole masters, -1 for cells no stencil reads)
  std::vector<int64_t> node_src;
  // edge ghost map, per edge slot 3*c+e (e = RT, DG, UP): source slot and sign
  std::vector<int64_t> edge_src;
  std::vector<int8_t> edge_sign;

  std::vector<double> X;         // coordinates, planes x | y | z, [3*ncells]
  std::vector<double> len;       // primal edge length, planes RT | DG | UP
  std::vector<double> pedlen;    // dual edge length, planes RT | DG | UP
  std::vector<double> area;      // hexagon part areas 1..6 (6 planes)
  std::vector<double> hex_inv;   // 1 / hexagon area
  std::vector<double> triarea;   // triangle area, planes LORT | UPLT
  std::vector<double> coriolis;  // f * dual area, planes LORT | UPLT
  std::vector<double> red_weight;  // visits of the reference's reduction sweep, folded onto owners

  // each lozenge's extended frame (cells [-EH, n+EH), EH = HALO + 1): node
  // coordinates and triangle circumcentres (LORT, UPLT), xyz interleaved;
  // kept for the inter-level geometry of the wavelet transforms
  int EH = 0, ES = 0;
  std::vector<double> xe, cle, cue;  // [10*ES*ES][3]
  // owned cell of every lattice point within WH cells of the owned square
  // (-1 where the lattice does not continue), frame [-WH, n+WH)^2
  int WH = 0, WS = 0;
  std::vector<int64_t> wide_src;
  size_t wcell(int lz, int i, int j) const {
    return (size_t)lz * WS * WS + (size_t)(j + WH) * WS + (size_t)(i + WH);
  }
  std::vector<double> mpe;           // edge midpoints [10*ES*ES][RT|DG|UP][3] (cells [-1, n+1])
  size_t ecell(int lz, int i, int j) const {
    return (size_t)lz * ES * ES + (size_t)(j + EH) * ES + (size_t)(i + EH);
  }
};

// Overlap areas of the scalar wavelet transform between a level and the next
// finer one (get_overl_areas, wavelet.f90:1408-1475, stored by set_RF_wgts
// 1512-1539): for every fine node that is the midpoint of a coarse edge,
// a(1..4) = areas of its dual hexagon shared with the dual cells of the
// edge's two end points and two opposite nodes.  Returns [4][fine ncells];
// zero at fine nodes that coincide with coarse nodes.
std::vector<double> overlap_areas(const Level& fine, const Level& coarse);

// min_level, domain_level: the reference's compile-time MIN_LEVEL and
// DOMAIN_LEVEL (src/param_J*.f90).  They fix how node coordinates are rounded
// (sub-domain corners first) and, through PATCH_SIZE = 2^(MIN_LEVEL -
// DOMAIN_LEVEL - 1), one ghost-list convention beside the south pole.
Level build_level(int level, int min_level, int domain_level, double radius, double omega);

// get_overl_areas for the child on edge e of coarse node (ip, jp) of lozenge
// lz: area[0..7] = a(1..4) and the split parts area(5..8), typ[0..1] =
// INSIDE 0 / OUTER1 1 / OUTER2 2 of the two sides.  False when the stencil
// leaves the extended frames.
bool overl_node(const Level& fine, const Level& coarse, int lz, int ip, int jp, int e, double area[8], int typ[2]);

// Inter-level weights of the velocity and flux transforms, computed on the
// host as the reference computes them at refinement (refine_patch.f90:
// 250-257, wavelet.f90:1176-1300, 1512-1705), with the 6 x 6 systems solved
// in LAPACK DGETF2/DGETRS order (dgesv.f90).
struct TransferWeights {
  int nc = 0;                  // coarse cells
  std::vector<double> split;   // [2][fine ncells]: overlap split parts a(5..6), halos as the owner's
  std::vector<double> iu;      // [9][fine ncells]: I_u_wgt of the outer child edges (owned fine cells)
  // R_F_wgt per coarse node and side: [4][6][coarse ncells], sets
  // 0 = RT k=0 (fine (2i+1, 2j-2), (2i+1, 2j-1)), 1 = RT k=1 ((2i, 2j), (2i, 2j+1)),
  // 2 = UP k=0 ((2i+1, 2j), (2i+1, 2j+1)), 3 = UP k=1 ((2i-2, 2j), (2i-2, 2j+1)),
  // each two triplets
  std::vector<double> rf;
};
TransferWeights transfer_weights(const Level& fine, const Level& coarse);

}  // namespace wtg
