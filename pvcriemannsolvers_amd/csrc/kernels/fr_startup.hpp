// Launch interface of the start-up limiters of nodal FR solutions
// (limiter/average.hpp, limiter/majority.hpp).
#ifndef MINICFD_KERNELS_FR_STARTUP_HPP_
#define MINICFD_KERNELS_FR_STARTUP_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcfd {

// u [cell][np][5] in place: every node takes the cell average with the
// global weights w_local[q] * det[cell * np + q] (det(J) at each node).
hipError_t LaunchFrAverage(double *u, const double *w_local, const double *det, int64_t n_cells, int np,
                           hipStream_t stream);
// value [cell][5] = the majority value of each component; *failed |= 1 where
// a component has none (value then 0).
hipError_t LaunchFrMajority(const double *u, double *value, int32_t *failed, int64_t n_cells, int np,
                            hipStream_t stream);
// u [cell][np][5] = value [cell][5] at every node.
hipError_t LaunchFrSetNodes(double *u, const double *value, int64_t n_cells, int np, hipStream_t stream);

}  // namespace mcfd

#endif  // MINICFD_KERNELS_FR_STARTUP_HPP_
