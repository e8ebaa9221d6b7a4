// Launch interface of the point-wise face-flux kernel (device solvers
// evaluated on given state pairs).
#ifndef MINICFD_KERNELS_FACE_FLUX_HPP_
#define MINICFD_KERNELS_FACE_FLUX_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcfd {

constexpr int kFluxUpwind = 0;
constexpr int kFluxInviscidWall = 1;
constexpr int kFluxSupersonicOutlet = 2;
constexpr int kFluxSubsonicInlet = 3;   // right = the given (p_total, u_cos, v_cos, w_cos, T_total)
constexpr int kFluxSubsonicOutlet = 4;  // right[4] = the back pressure

// flux[i] of point i (device arrays): frames [n][9] rows (n, t, b), left and
// right [n][5] conservative, star_state [n] the exact solver's per-point state
// (in/out, or null for fresh solvers).  form 0: the FR stage kernels' code;
// form 1: the modal-DG face pass's code.  kind: kFlux*.
hipError_t LaunchFaceFluxPoints(int riemann, int form, int kind, int64_t n, const double *frames,
                                const double *left, const double *right, uint8_t *star_state, double *flux,
                                hipStream_t stream);

}  // namespace mcfd

#endif  // MINICFD_KERNELS_FACE_FLUX_HPP_
