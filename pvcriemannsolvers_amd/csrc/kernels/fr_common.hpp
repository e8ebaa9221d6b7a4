// Device helpers of the FR stage kernels on affine hexahedra: line indexing
// of faces, the fixed-order extrapolation to a line end, the XCD-aware block
// order, the columns of the local flux F~ = F J*, the lifted flux gaps and
// the RK stage combination.
#ifndef MINICFD_KERNELS_FR_COMMON_HPP_
#define MINICFD_KERNELS_FR_COMMON_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "euler.hpp"
#include "fr_hex.hpp"

namespace mcfd {
namespace fr {

// Line of solution points through in-face point (a, b) of local face f:
// node(m) = a * A + b * B + m * S.  Branch-free for lane-varying f.
struct LineIndex { int a_mul, b_mul, stride, side, dir; };
__device__ __forceinline__ LineIndex FaceLineIndex(int q, int f) {
  const int dir = (2118 >> (2 * f)) & 3;  // {2, 1, 0, 1, 0, 2}
  const int side = (0x2C >> f) & 1;       // {0, 0, 1, 1, 0, 1}
  LineIndex li;
  li.dir = dir;
  li.side = side;
  li.a_mul = dir == 0 ? q : q * q;
  li.b_mul = dir == 2 ? q : 1;
  li.stride = dir == 0 ? q * q : (dir == 1 ? q : 1);
  return li;
}
// The same line in the skewed LDS layout of Shape<Q>.
template <int Q>
__device__ __forceinline__ LineIndex SkewedLineIndex(int f) {
  constexpr int SI = Q * (Q + 1), SJ = Q + 1;
  const int dir = (2118 >> (2 * f)) & 3;
  LineIndex li;
  li.dir = dir;
  li.side = (0x2C >> f) & 1;
  li.a_mul = dir == 0 ? SJ : SI;
  li.b_mul = dir == 2 ? SJ : 1;
  li.stride = dir == 0 ? SI : (dir == 1 ? SJ : 1);
  return li;
}

// Extrapolation to a line end: a fixed-order fused chain so that both cells of
// a face evaluate a given trace with identical rounding.
template <int Q>
__device__ __forceinline__ double Extrapolate(const double *w, const double *v) {
  double acc = w[0] * v[0];
#pragma unroll
  for (int m = 1; m < Q; ++m) acc = __fma_rn(w[m], v[m], acc);
  return acc;
}

__device__ __forceinline__ int64_t LogicalBlock(int32_t remap) {
  int64_t b = blockIdx.x;
  if (!remap) return b;
  // Blocks are dispatched round-robin over the 8 XCDs; give each XCD one
  // contiguous range of cells so neighbour traces hit its own L2.
  int64_t n = gridDim.x, per = n / 8, rem = n % 8;
  int64_t xcd = b % 8, idx = b / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Column d of the local flux F~ = F J* at one solution point.
template <bool kDiag>
__device__ __forceinline__ void LocalFluxColumn(State const &c, State const &p, const double *js, int d,
                                                double *out) {
  const double rh = c.v[4] + p.v[4];
  if constexpr (kDiag) {
    const double vel = p.v[1 + d], s = js[4 * d];
    out[0] = c.v[1 + d] * s;
    out[1] = (c.v[1] * vel + (d == 0 ? p.v[4] : 0.0)) * s;
    out[2] = (c.v[2] * vel + (d == 1 ? p.v[4] : 0.0)) * s;
    out[3] = (c.v[3] * vel + (d == 2 ? p.v[4] : 0.0)) * s;
    out[4] = rh * vel * s;
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) out[k] = 0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const double vel = p.v[1 + e], s = js[3 * e + d];
      out[0] += c.v[1 + e] * s;
      out[1] += (c.v[1] * vel + (e == 0 ? p.v[4] : 0.0)) * s;
      out[2] += (c.v[2] * vel + (e == 1 ? p.v[4] : 0.0)) * s;
      out[3] += (c.v[3] * vel + (e == 2 ? p.v[4] : 0.0)) * s;
      out[4] += rh * vel * s;
    }
  }
}

// Lifted flux gaps at solution point (ni, nj, nk) (fr/general.hpp:409-434):
// the six faces' gaps times the Vincent derivatives g', summed in a fixed
// order.  g holds one component's gaps face-major, g[f * Q^2 + a * Q + b].
template <int Q>
__device__ __forceinline__ double LiftedGaps(const double *g, const LineTables &lt, int ni, int nj, int nk) {
  constexpr int NFP = Q * Q;
  double c = g[5 * NFP + ni * Q + nj] * lt.gp[1][nk];
  c = __fma_rn(g[0 * NFP + ni * Q + nj], lt.gp[0][nk], c);
  c = __fma_rn(g[1 * NFP + ni * Q + nk], lt.gp[0][nj], c);
  c = __fma_rn(g[3 * NFP + ni * Q + nk], lt.gp[1][nj], c);
  c = __fma_rn(g[4 * NFP + nj * Q + nk], lt.gp[0][ni], c);
  return __fma_rn(g[2 * NFP + nj * Q + nk], lt.gp[1][ni], c);
}

// Residual -> RK stage output (temporal/rk.hpp:31-49 with ode.hpp:121-128):
// stage 0 returns the residual; 1: us + dt r; 2: (3 u0 + (us + dt r)) / 4;
// 3: (2 (us + dt r) + u0) / 3, each operation rounded on its own.
__device__ __forceinline__ double RkOutput(int stage, double res, double us, double u0, double dt) {
  if (stage == 0) return res;
  const double next = AddRn(us, MulRn(res, dt));
  if (stage == 1) return next;
  if (stage == 2) return AddRn(MulRn(u0, 3.0), next) / 4.0;
  return AddRn(MulRn(next, 2.0), u0) / 3.0;
}

}  // namespace fr
}  // namespace mcfd

#endif  // MINICFD_KERNELS_FR_COMMON_HPP_
