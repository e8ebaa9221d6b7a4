// One mesh discretised by modal DG (spatial/dg/general.hpp) on affine
// hexahedra or tetrahedra: host connectivity, per-cell orthonormal bases and
// geometry, face records, reference-element tables and the device arrays the
// kernels in dg_modal.hip read.
#ifndef MINICFD_HOST_DG_GROUP_HPP_
#define MINICFD_HOST_DG_GROUP_HPP_

#include <array>
#include <cstdint>
#include <vector>

#include "../kernels/dg_modal.hpp"
#include "common.hpp"

namespace mcfd {

struct DgFace {
  int64_t holder, sharer;  // sharer < 0: boundary
  int holder_face, sharer_face;
  int orient;              // holder-sorted face node i sits at the sharer's sorted position Perm(orient, i)
  int tag;                 // boundary tag (interior: -1)
};

struct DgGroup {
  int cell_type = 0, degree = 0, riemann = 0;
  int n = 0, nq = 0, nf = 0, nfp = 0;  // basis functions, volume points, faces and points per face
  int64_t n_cells = 0, n_interior = 0;
  std::vector<DgFace> faces;           // interior faces first, then boundary faces
  DgTables tables{};
  std::vector<DgTables> h_tables;     // upload source of `tables`
  std::vector<double> h_basis, h_geo, h_faces, h_bc_state;
  std::vector<int32_t> h_nbr, h_finfo, h_bc_kind;
  std::vector<double> qp;              // global quadrature points [cell][q][3]
  DeviceArray<double> u0, u1, u2, res, basis, geo, face_rec, bc_state;
  DeviceArray<double> ct0, ct1, ct2;  // Taylor coefficients of u0, u1, u2 (kept current)
  DeviceArray<int32_t> finfo, bc_kind;
  // the stage's face pass: w Fhat once per face point, then the cell pass
  std::vector<int32_t> h_face_cells;  // [face][3]: holder, sharer or -(1 + tag), faces and orientation
  DeviceArray<int32_t> face_cells;
  DeviceArray<double> fhat;           // [face][point][5]
  DeviceArray<uint8_t> exact_star;    // exact solver: [face][point] sign of the last star velocity
  DeviceArray<DgTables> tables_d;
  // WENO limiter (MCFD_LIMITER_*), applied to every state that is set
  int limiter = 0;
  double weno_w0 = 0.001, weno_eps = 1e-6;
  std::vector<int32_t> h_adj, h_adj_face, h_nadj;  // Cell::adj_cells_/adj_faces_ in face order, [cell][6]
  DeviceArray<int32_t> adj, adj_face, nadj;
  DeviceArray<uint8_t> flags;
  DeviceArray<double> flag_rec;        // [cell][kDgFlagRecord], the detector's view of the current state
  DeviceArray<int32_t> troubled;       // flagged cells of the last limiter pass, then their count
  DeviceArray<double> proj, eig;

  void Build(const mcfd_mesh_desc &mesh, const mcfd_scheme_desc &scheme,
             const std::vector<int32_t> &bc_kinds, const std::vector<double> &bc_states);
  void Upload(hipStream_t stream);
  // polynomial().LocalToGlobalAndValue of every cell at n_local local points
  // (hexahedra: [-1, 1]^3; tetrahedra: the corners (1,0,0), (0,1,0), (0,0,1),
  // (0,0,0) of CGNS vertices 0..3): xyz [cell][p][3], values [cell][p][5] of
  // the modal column; either output may be null
  void Evaluate(const double *column, int n_local, const double *local, double *xyz, double *values) const;

  size_t ColumnSize() const { return size_t(n_cells) * n * 5; }
  // Stage 0 residual, 1..3 SSP-RK3 stages (Taylor coefficients, then the stage kernel).
  // With dt_min_bits set (no limiter only), the stage also folds the time-step
  // reduction of its output state into it.
  void Stage(int stage, double dt, hipStream_t stream, unsigned long long *dt_min_bits = nullptr);
  // Limiter pass over the state `u` with Taylor coefficients `ct`.
  void Limit(double *u, double *ct, hipStream_t stream);
  // The state u0 was set from outside: form its Taylor coefficients (and the
  // detector record) and, unless `limit` is false (a restored state), limit it.
  void SolutionChanged(hipStream_t stream, bool limit = true);
  // min over cells of length / max wave speed at the first N quadrature points.
  void TimeStepBits(unsigned long long *dt_min_bits, hipStream_t stream);
  // u0 = coefficients of f given at the quadrature points ([cell][nq][5] host
  // array, polynomial/projection.hpp:28-43), projected on the device; then
  // SolutionChanged.
  void Approximate(const double *values, hipStream_t stream);
  void Faces(int64_t *holder, int64_t *sharer, int32_t *holder_face, int32_t *sharer_face) const;
  // Tag and quadrature points of every boundary face ([boundary face][point][3]).
  void BoundaryPoints(std::vector<int32_t> *tags, std::vector<double> *xyz) const;
  // Given states per boundary face point (BoundaryTable), or null: the tags' constants.
  const double *bc_point_state = nullptr;
};

}  // namespace mcfd

#endif  // MINICFD_HOST_DG_GROUP_HPP_
