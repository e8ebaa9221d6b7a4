// One mesh discretised by nodal FR on hexahedra: host connectivity, line
// tables, geometry/frame records and the device arrays the stage kernel
// reads.  A solver handle owns one group; a batch owns one per degree.
// Meshes of affine cells keep one deduplicated geometry record per cell and
// one frame per face; a mesh with a trilinear cell ("curved") keeps the
// metric of every solution point, the signed scale of every cell flux point
// and the frame of every face flux point.
#ifndef MINICFD_HOST_FR_GROUP_HPP_
#define MINICFD_HOST_FR_GROUP_HPP_

#include <array>
#include <cstdint>
#include <vector>

#include "../kernels/fr_hex.hpp"
#include "common.hpp"
#include "hex_connectivity.hpp"

namespace mcfd {

struct FrHexGroup {
  int q = 0, np = 0, degree = 0, riemann = 0;
  bool mixed = false, diag_metric = false, curved = false;
  int64_t n_cells = 0;
  HexConnectivity conn;
  LineTables lt;
  std::vector<double> x1d, w1d;  // 1-D solution points and their quadrature weights
  std::vector<std::array<double, 3>> corner0;  // node 0 of each cell
  // host copies uploaded by Upload()
  std::vector<int32_t> h_nbr, h_finfo, h_geo_id, h_bc_kind, h_cell_case, h_tile_order;
  std::vector<double> h_geo, h_frames, h_bc_state;
  // curved meshes: [cell][node][kNodeGeoStride] and [cell][6][Q * Q]
  std::vector<double> h_geo_node, h_fscale;
  DeviceArray<double> geo_node, fscale;
  DeviceArray<double> u0, u1, u2, res;
  DeviceArray<int32_t> nbr, finfo, geo_id, bc_kind, cell_case, tile_order;
  DeviceArray<double> geo, frames, bc_state, case_dt;
  // exact solver: per (cell, local face, flux point) state of its solver
  // object, the sign of the last star velocity (zero at creation)
  DeviceArray<uint8_t> exact_star;
  // boundary face index of each boundary cell face, [cell][6] (-1 elsewhere),
  // for the per-point boundary values of BoundaryTable
  std::vector<int32_t> h_bface_of;
  DeviceArray<int32_t> bface_of;

  // Host-side construction; bc_kind/bc_state index the mesh's boundary tags.
  void Build(const mcfd_mesh_desc &mesh, const mcfd_scheme_desc &scheme,
             const std::vector<int32_t> &bc_kinds, const std::vector<double> &bc_states);
  void BuildCurved(const mcfd_scheme_desc &scheme);
  // Order in which the stage kernel sweeps its tiles of `tile_cells`
  // consecutive cells: Morton order of the tile centres, so that the tiles in
  // flight on one XCD are spatial neighbours and their face-trace gathers hit
  // that XCD's L2.  Uploaded by Upload(); without it tiles go in id order.
  void BuildTileOrder(int tile_cells);
  // Device allocation and upload on `stream` (not synchronised).
  void Upload(hipStream_t stream);

  size_t ColumnSize() const { return size_t(n_cells) * np * 5; }
  // Kernel arguments of one pass: stage 0 residual, 1..3 SSP-RK3 stages.
  FrHexArgs Args(int stage, double dt) const;
  void SamplePoints(double *xyz) const;
  // polynomial().LocalToGlobalAndValue of every cell at n_local local points
  // (xi, eta, zeta in [-1, 1]^3): xyz [cell][p][3], values [cell][p][5] of
  // the nodal column; either output may be null
  void Evaluate(const double *column, int n_local, const double *local, double *xyz, double *values) const;
  // det(J) at every solution point, [cell][node] (the start-up average's weights)
  std::vector<double> NodeDeterminants() const;
  // Tag and flux points of every boundary face ([boundary face][a * Q + b][3],
  // (a, b) the in-face indices of the holder's local face).
  void BoundaryPoints(std::vector<int32_t> *tags, std::vector<double> *xyz) const;
  void Faces(int64_t *holder, int64_t *sharer, int32_t *holder_face, int32_t *sharer_face,
             int32_t *fp_ijk) const;
};

}  // namespace mcfd

#endif  // MINICFD_HOST_FR_GROUP_HPP_
