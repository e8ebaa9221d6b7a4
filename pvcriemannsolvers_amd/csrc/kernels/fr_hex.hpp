// Launch interface of the nodal FR kernels on hexahedra (affine or trilinear).
#ifndef MINICFD_KERNELS_FR_HEX_HPP_
#define MINICFD_KERNELS_FR_HEX_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcfd {

constexpr int kMaxQ = 6;
constexpr int kMaxBatchGroups = 4;

// 1-D tables of the solution-point line (Q points):
//   d[i][m]   = l_m'(x_i)          (Lagrange derivative matrix)
//   lend[e][m] = l_m(-1), l_m(+1)   (extrapolation to the line ends)
//   gp[e][m]  = g_L'(x_m), g_R'(x_m) (Vincent correction derivatives)
struct LineTables {
  double d[kMaxQ][kMaxQ];
  double lend[2][kMaxQ];
  double gp[2][kMaxQ];
};

// Per-cell affine geometry record (20 doubles):
//   [0..8]  J* = det(J) J^{-1} (row-major, F~ = F J*)
//   [9]     1 / det(J)
//   [10..15] signed face scale +-|J*.col(normal dir)| of local faces 0..5
//   [16]    length = cbrt(volume) / 2
constexpr int kGeoStride = 20;
constexpr int kFrameStride = 6;  // n, t
// Per-node record of a trilinear cell: J* (9, as above), 1 / det(J).
constexpr int kNodeGeoStride = 10;

struct FrHexArgs {
  const double *us;      // state of this stage, also read for neighbour traces
  const double *u0;      // U0 for RK stages 2 and 3
  double *out;           // residual (stage 0) or next state
  const int32_t *nbr;    // [cell][6]: neighbour cell, or -(1 + boundary tag)
  const int32_t *finfo;  // [cell][6]: frame id << 7 | holder << 6 | orient << 3 | neighbour face
  const int32_t *geo_id; // [cell]
  const double *geo;     // [geo][kGeoStride]
  const double *frames;  // [frame][kFrameStride]; curved: [face][a * Q + b][kFrameStride]
  // curved (trilinear) meshes: [cell][node][kNodeGeoStride] and the signed
  // flux-point scales [cell][local face][a * Q + b]; null otherwise
  const double *geo_node;
  const double *fscale;
  const int32_t *bc_kind;   // [tag]: BcKind (euler.hpp)
  const double *bc_state;   // [tag][5] given state (conservative, or the subsonic parameters)
  // If set, the given state of every boundary flux point instead:
  // [boundary face][a * Q + b][5], with bface_of [cell][6] the boundary face
  // index of each boundary cell face (the tag's boundary function at the stage time)
  const double *bc_point_state;
  const int32_t *bface_of;
  const int32_t *cell_case; // [cell] case index, or null (one time step for all cells)
  const double *case_dt;    // [case] time step when cell_case is set
  int64_t n_cells;
  double dt;
  int32_t stage;         // 0 residual, 1..3 SSP-RK3 stage
  int32_t xcd_remap;     // remap block ids so each XCD sweeps a contiguous range
  int32_t diag_metric;   // every J* is diagonal (axis-aligned cells)
  int32_t mixed;         // FP32 flux-column contractions (mixed-precision form)
  int32_t curved;        // per-node metrics and per-point scales and frames
  const int32_t *tile_order;  // [tile]: tile swept by logical block b (spatial locality), or null
  // If set, the time-step reduction of LaunchFrHexTimeStep over the state this
  // pass writes is folded into *dt_min_bits (same arithmetic, no extra pass).
  unsigned long long *dt_min_bits;
  // exact solver only: [cell][local face][point] sign of each flux point's
  // last star velocity (ExactFlux), zero-initialised at creation
  uint8_t *exact_star;
};

// Launches one residual / RK-stage pass for Q points per direction and the
// given Riemann solver on `stream`.  Returns hipSuccess or the launch error.
hipError_t LaunchFrHexStage(int q, int riemann, const FrHexArgs &args, const LineTables &lt,
                            hipStream_t stream);
// Cells per workgroup (one tile) of LaunchFrHexStage for Q points per direction.
int FrHexStageTileCells(int q);

// One degree group of a batched launch: its own arrays, line tables and tile
// variant (cells per workgroup, see FrHexTileCells).
struct FrHexBatchGroup {
  FrHexArgs args;
  LineTables lt;
  int32_t q;
  int32_t tile;
  int64_t first_item;   // first work item of this group
};
struct FrHexBatchArgs {
  const FrHexBatchGroup *group;  // device array of n_groups entries (one stage)
  int32_t n_groups;
  int32_t max_tile_lds;          // LDS bytes of the largest tile in use
  int64_t n_items;
};

// Cells per workgroup and LDS bytes of tile variant `tile` (0..2) for q
// points (2..4); 0 if invalid.
int FrHexTileCells(int q, int tile);
int FrHexTileLdsBytes(int q, int tile);

// Persistent batched stage over all groups with n_workgroups workgroups.
hipError_t LaunchFrHexBatch(int riemann, bool mixed, const FrHexBatchArgs &args, int n_workgroups,
                            hipStream_t stream);
// Workgroups of the batched kernel resident on the whole device with the
// given dynamic LDS (0 on failure).
int FrHexBatchResidentWorkgroups(int riemann, bool mixed, int lds_bytes);

// Per-cell minimum of length / max wave speed over the solution points
// (spatial/fem.hpp:417-435), reduced to one double in *dt_min (device).
// *bits = the ordered bits of 1e300, the start of a time-step reduction.
hipError_t LaunchResetTimeStepBits(unsigned long long *bits, hipStream_t stream);
hipError_t LaunchFrHexTimeStep(int q, const double *u, const int32_t *geo_id, const double *geo,
                               int64_t n_cells, unsigned long long *dt_min_bits, hipStream_t stream);

}  // namespace mcfd

#endif  // MINICFD_KERNELS_FR_HEX_HPP_
