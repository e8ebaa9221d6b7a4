// Host-side connectivity of conforming hexahedral meshes for the nodal FR
// kernels: faces, holder/sharer, local face ids, sharer orientation, normal
// frames and per-cell geometry (affine or trilinear).
//
// Semantics follow the reference's face construction:
//   * interior faces: one per adjacency pair (holder = lower cell id), in the
//     order (holder, sharer) ascending (mesh/part.hpp:932-975 over the pairs
//     written by mesh/shuffler.hpp with i < j);
//   * face node order: the holder's CGNS face table, so the normal points
//     from the holder to the sharer (coordinate/hexahedron.hpp:84-116,
//     379-389, coordinate/cell.hpp:235-258);
//   * boundary faces follow the interior ones in input order
//     (mesh/part.hpp:1456-1549);
//   * frame = {n, t, b}: n = normalize(J_xi x J_eta), t = normalize(J_xi),
//     b = n x t of the face map (coordinate/face.hpp:54-62).
// The construction itself is this library's own: faces are matched by
// sorted node-id keys, periodic faces by translated centres, and the sharer
// orientation is read off three corner correspondences.
#ifndef MINICFD_HOST_HEX_CONNECTIVITY_HPP_
#define MINICFD_HOST_HEX_CONNECTIVITY_HPP_

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/minicfd.h"

namespace mcfd {

// In-face node indices (a, b) of local face f in terms of the cell's (i, j, k):
//   f = 0 (zeta=-1), 5 (zeta=+1): (i, j);  f = 1, 3 (eta=-/+1): (i, k);
//   f = 2, 4 (xi=+/-1): (j, k).
constexpr int kHexFaceNormalDir[6] = { 2, 1, 0, 1, 0, 2 };
constexpr int kHexFaceSide[6] = { 0, 0, 1, 1, 0, 1 };  // 0: -1 end, 1: +1 end

struct HexFace {
  int64_t holder, sharer;  // sharer < 0: boundary
  int8_t holder_face, sharer_face;
  // sharer in-face indices from the holder's: bit0 swap, bit1 flip a, bit2 flip b
  int8_t orient;
  // face quadrature point (i, j) -> holder in-face (a, b), same encoding
  int8_t quad_to_holder;
  int32_t tag;             // boundary tag (interior: -1)
  double frame[6];         // n, t (b = n x t) at the face centre
  double shift[3];         // sharer coordinates = holder coordinates + shift
  double corners[12];      // the holder's face corners in its CGNS face order (the face map)
};

struct HexConnectivity {
  int64_t n_cells = 0;
  std::vector<HexFace> faces;  // interior first, then boundary
  int64_t n_interior = 0;
  // per cell: affine map x = x0 + J^T (xi + 1) / 2 ... summarised as
  // J (rows d/dxi, d/deta, d/dzeta), det(J) and the centre.
  std::vector<std::array<double, 9>> jac;
  std::vector<double> det;
  // per cell: the 8 CGNS corners and whether the map is affine (a
  // parallelepiped); a trilinear cell keeps its centre values in jac / det
  std::vector<std::array<double, 24>> vertices;
  std::vector<uint8_t> affine;
};

// Trilinear (Hexahedron8) map of corners xv[8][3] at local point xi[3], and its
// Jacobian j (rows d/dxi, d/deta, d/dzeta), coordinate/hexahedron.hpp:397-526.
void TrilinearPoint(const double *xv, const double *xi, double *x);
void TrilinearJacobian(const double *xv, const double *xi, double *j);
// Normal frame {n, t} of the bilinear face map over corners xq[4][3] (CGNS
// QUAD_4 order) at local (a, b): n = normalize(J_a x J_b), t = normalize(J_a)
// (coordinate/face.hpp:54-62).
void FaceFrameAt(const double *xq, double a, double b, double *frame);

// Maps in-face indices (a, b) through an orientation code for q points.
inline void ApplyOrient(int code, int q, int a, int b, int *a2, int *b2) {
  int s = (code & 1) ? b : a, t = (code & 1) ? a : b;
  *a2 = (code & 2) ? q - 1 - s : s;
  *b2 = (code & 4) ? q - 1 - t : t;
}

// Node index on the line of local face f through in-face (a, b), position m
// counted from the xi/eta/zeta = -1 end.
inline int HexLineNode(int q, int f, int a, int b, int m) {
  switch (f) {
  case 0: case 5: return (a * q + b) * q + m;
  case 1: case 3: return (a * q + m) * q + b;
  default: return (m * q + a) * q + b;
  }
}

// Builds the connectivity; throws std::invalid_argument on bad input.
HexConnectivity BuildHexConnectivity(const mcfd_mesh_desc &mesh);

}  // namespace mcfd

#endif  // MINICFD_HOST_HEX_CONNECTIVITY_HPP_
