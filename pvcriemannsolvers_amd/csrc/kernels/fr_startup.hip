// Start-up limiters of nodal solutions (FR on hexahedra), one thread per
// (cell, component):
//   average  limiter/average.hpp:11-28  every node takes the cell's average
//            sum_q u_q w_q / sum_q w_q over the solution points, with the
//            global weights w_q of polynomial/hexahedron.hpp:626-632;
//   majority limiter/majority.hpp:14-42  every node takes the value that
//            more than half of the cell's nodes hold; scanning the nodes in
//            order, the first whose running count exceeds N / 2 decides.
#include "fr_startup.hpp"

#include "euler.hpp"

namespace mcfd {

namespace {

__global__ __launch_bounds__(256) void FrAverageKernel(double *u, const double *w_local, const double *det,
                                                       int64_t n_cells, int np) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n_cells * 5) return;
  const int64_t cell = t / 5;
  const int k = (int)(t - cell * 5);
  double *uc = u + cell * np * 5;
  double sum = 0, volume = 0;
  for (int q = 0; q < np; ++q) {
    const double w = MulRn(w_local[q], det[cell * np + q]);
    sum = AddRn(sum, MulRn(uc[q * 5 + k], w));
    volume = AddRn(volume, w);
  }
  const double average = sum / volume;
  for (int q = 0; q < np; ++q) uc[q * 5 + k] = average;
}

__global__ __launch_bounds__(256) void FrMajorityKernel(const double *u, double *value, int32_t *failed,
                                                        int64_t n_cells, int np) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n_cells * 5) return;
  const int64_t cell = t / 5;
  const int k = (int)(t - cell * 5);
  const double *uc = u + cell * np * 5;
  bool found = false;
  double v = 0;
  for (int q = 0; q < np && !found; ++q) {
    const double x = uc[q * 5 + k];
    int count = 0;
    for (int r = 0; r <= q; ++r) count += uc[r * 5 + k] == x;
    if (count * 2 > np) {
      found = true;
      v = x;
    }
  }
  value[t] = v;
  if (!found) atomicOr(failed, 1);
}

__global__ __launch_bounds__(256) void FrSetNodesKernel(double *u, const double *value, int64_t n_cells, int np) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n_cells * np * 5) return;
  const int64_t cell = t / (np * 5);
  const int k = (int)(t % 5);
  u[t] = value[cell * 5 + k];
}

dim3 Blocks(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t LaunchFrAverage(double *u, const double *w_local, const double *det, int64_t n_cells, int np,
                           hipStream_t stream) {
  if (n_cells < 1) return hipSuccess;
  hipLaunchKernelGGL(FrAverageKernel, Blocks(n_cells * 5), dim3(256), 0, stream, u, w_local, det, n_cells, np);
  return hipGetLastError();
}

hipError_t LaunchFrMajority(const double *u, double *value, int32_t *failed, int64_t n_cells, int np,
                            hipStream_t stream) {
  if (n_cells < 1) return hipSuccess;
  hipLaunchKernelGGL(FrMajorityKernel, Blocks(n_cells * 5), dim3(256), 0, stream, u, value, failed, n_cells, np);
  return hipGetLastError();
}

hipError_t LaunchFrSetNodes(double *u, const double *value, int64_t n_cells, int np, hipStream_t stream) {
  if (n_cells < 1) return hipSuccess;
  hipLaunchKernelGGL(FrSetNodesKernel, Blocks(n_cells * np * 5), dim3(256), 0, stream, u, value, n_cells, np);
  return hipGetLastError();
}

}  // namespace mcfd
