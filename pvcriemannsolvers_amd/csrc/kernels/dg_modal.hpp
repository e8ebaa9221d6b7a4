// Modal DG on affine hexahedra and tetrahedra (spatial/dg/general.hpp:51-202):
// device tables, kernel arguments and launchers.
//
// Every cell carries the orthonormal basis phi = S T(x - c) of the reference
// (basis/linear.hpp:87-141): T the Taylor monomials of degree p <= 2, S the
// lower-triangular Gram-Schmidt factor.  The kernels work in the Taylor
// space: a prep pass forms the Taylor coefficients Ct = C S of every cell
// (so u(x) = Ct T(x - c)), the stage pass evaluates states at the volume and
// face points from Ct, accumulates the residual's Taylor moments
//   Rt[k][j] = sum_q w_q F(u_q)[k] . grad T_j(y_q) - sum_f w_f Fhat_f[k] T_j(y_f)
// and maps them back with R = Rt S^T.  Points are kept in the cell's local
// coordinates; y = J (xi - xi_c) is the offset from the cell centre.
#ifndef MINICFD_KERNELS_DG_MODAL_HPP_
#define MINICFD_KERNELS_DG_MODAL_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcfd {

constexpr int kDgMaxVolumePoints = 27;  // 3^3 Gauss points (hexahedra, p = 2)
constexpr int kDgMaxFacePoints = 9;     // 3^2 Gauss points (quadrangles, p = 2)
constexpr int kDgMaxOrient = 8;         // dihedral group of the square
constexpr int kDgGeoStride = 16;        // J rows (9), |det J| (1), length (1), centre (3), volume (1), pad
constexpr int kDgFaceStride = 8;        // n (3), t (3), area factor (1), pad
// Per-cell record that the WENO troubled-cell detector reads (two 128-byte
// lines instead of the scattered coefficient, state, basis and geometry
// lines): rho Taylor coefficients [0, N), E Taylor coefficients [10, 10 + N),
// centre [20, 23), volume [23], rho and E average terms u[0][k] * S[0][0]
// [24], [25].  The centre and volume are set once; the Taylor kernel and the
// stage pass rewrite the rest for every state they produce.
constexpr int kDgFlagRecord = 32;
constexpr int kDgFlagCentre = 20, kDgFlagVolume = 23, kDgFlagAverage = 24;
// Set once on the host with the C library's pow: the detector's volume power
// volume^((p + 1) / 6) (weno.hpp:252) and the smoothness scales
// (1 / S00^2)^(2/3 - 1), (1 / S00^2)^(4/3 - 1) (weno.hpp:74-95).
constexpr int kDgFlagDivisor = 26, kDgFlagScale1 = 27, kDgFlagScale2 = 28;

// Reference-element tables of one (cell type, degree); lives in device memory.
struct DgTables {
  double vol_eta[kDgMaxVolumePoints][3];  // xi_q - xi_c
  double vol_w[kDgMaxVolumePoints];       // local weights
  // face point q of a face whose holder-sorted node i sits at the cell's
  // sorted face position Perm(orient, i): its local coordinates minus xi_c
  double face_eta[6][kDgMaxOrient][kDgMaxFacePoints][3];
  double face_w[kDgMaxFacePoints];        // local weights on the face
};

struct DgModalArgs {
  const double *us;       // state of this stage [cell][N][5]
  const double *u0;       // U0 for RK stages 2 and 3
  double *out;            // residual (stage 0) or next state
  const double *ct;       // Taylor coefficients of `us`, [cell][N][5]
  double *ct_out;         // Taylor coefficients of `out`, or null
  const double *basis;    // [cell][N(N+1)/2] S, packed lower triangle by rows
  const double *geo;      // [cell][kDgGeoStride]
  const int32_t *finfo;   // [cell][6]: face id << 8 | holder << 7 | orient << 3 | neighbour's local face
  const double *faces;    // [face][kDgFaceStride]
  const int32_t *bc_kind; // [tag]: BcKind (euler.hpp)
  const double *bc_state; // [tag][5] given state (conservative, or the subsonic parameters)
  // If set, the given state of every boundary face point instead,
  // [boundary face][point][5] (the tag's boundary function at the stage time)
  const double *bc_point_state;
  int64_t n_interior;     // interior faces, which come first
  const DgTables *tables;
  int64_t n_cells;
  double dt;
  int32_t stage;          // 0 residual, 1..3 SSP-RK3 stage
  // If set (with ct_out), the time-step reduction of LaunchDgTimeStep over the
  // state this pass writes is folded into *dt_min_bits (same arithmetic).
  unsigned long long *dt_min_bits;
  double *flag_rec;       // [cell][kDgFlagRecord] for the WENO detector, or null (no limiter)
  // The pass runs as two launches: w Fhat once per face point into fhat
  // [face][point][5], then the cells read it.  face_cells [face][3]: holder,
  // sharer or -(1 + boundary tag), holder face | sharer face << 3 |
  // orientation << 6.
  double *fhat;
  const int32_t *face_cells;
  int64_t n_faces;
  // exact solver only: [face][point] sign of each face point's last star
  // velocity (ExactFlux), zero-initialised at creation, updated every pass
  uint8_t *exact_star;
};

// Ct = C S for every cell of `us` into `ct`.
hipError_t LaunchDgTaylorCoefficients(int degree, const double *us, const double *basis, int64_t n_cells,
                                      double *ct, double *flag_rec, hipStream_t stream);
// One residual / RK-stage pass (reads a.ct, which must hold the Taylor
// coefficients of a.us, and writes those of a.out to a.ct_out if set).
hipError_t LaunchDgModalStage(int cell_type, int degree, int riemann, const DgModalArgs &a, hipStream_t stream);
// WENO limiter (limiter/weno.hpp, limiter/reconstruct.hpp) over a state
// column `u` whose Taylor coefficients are in `ct`.
struct DgLimiterArgs {
  double *u;              // state, reconstructed in place for troubled cells
  double *ct;             // Taylor coefficients of u, updated for troubled cells
  const double *basis;    // as DgModalArgs
  const double *geo;
  const int32_t *adj;     // [cell][6] neighbours in interior-face order (Cell::adj_cells_)
  const int32_t *adj_face;  // [cell][6] the faces to them (Cell::adj_faces_)
  const int32_t *n_adj;   // [cell]
  const double *faces;    // face records, as DgModalArgs
  const DgTables *tables;
  const double *flag_rec; // [cell][kDgFlagRecord] of the state u
  uint8_t *flags;         // [cell]
  int32_t *troubled;      // [cell] list of the flagged cells, in no particular order
  int32_t *n_troubled;    // its length
  double *proj;           // scratch, n_cells * 7 * 5 N doubles.  weno::Eigen: [cell][7][N][5] neighbour
                          // projections, then the cell's own.  weno::Lazy: [troubled-list position][2][N][5]
                          // the reconstructed state and its Taylor coefficients, until written back
  double *eig;            // scratch [cell][6][5][N]: blended characteristic fields per face (Eigen)
  int64_t n_cells;
  int32_t mode;           // MCFD_LIMITER_*: 1, 2 weno::Lazy on flagged / all cells; 3, 4 weno::Eigen
  double w0, eps;
};
hipError_t LaunchDgLimiter(int cell_type, int degree, const DgLimiterArgs &a, hipStream_t stream);

// Modal coefficients [cell][N][5] of point values [cell][nq][5] given at the
// cell quadrature points (projection.hpp:28-43).
hipError_t LaunchDgApproximate(int degree, const double *values, const double *basis, const double *geo,
                               const DgTables *tables, int nq, int64_t n_cells, double *coeffs, hipStream_t stream);

// min over cells of length / max_{q < N} (|u| + c)(x_q), as ordered bits of a
// positive double folded into *dt_min_bits with atomicMin.
hipError_t LaunchDgTimeStep(int cell_type, int degree, const double *ct, const double *geo,
                            const DgTables *tables, int64_t n_cells, unsigned long long *dt_min_bits,
                            hipStream_t stream);

}  // namespace mcfd

#endif  // MINICFD_KERNELS_DG_MODAL_HPP_
