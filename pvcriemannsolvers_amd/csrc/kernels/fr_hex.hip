// Nodal flux reconstruction on affine hexahedra for CDNA4 (gfx950), FP64.
//
// One pass = residual of spatial/fr/general.hpp (volume term :271-288, face
// correction :296-315,409-434, boundary corrections :435-640, det(J) scaling
// :249-268) fused with one SSP-RK3 stage of temporal/rk.hpp:31-49.
//
// Work decomposition: a 256-thread workgroup owns C consecutive cells
// (C * Q^3 <= 256 solution points).  Each pass
//   0. streams the cells' states (contiguous in the reference column layout)
//      into LDS with coalesced flat loads;
//   1. computes the local flux F~ = F J* per solution point into LDS;
//   2. computes, per face point of its own 6 faces, the own trace (LDS), the
//      neighbour trace (global loads through L2 -- the neighbour's line of Q
//      points), the upwind flux and the flux gap; each interior face is thus
//      solved by both of its cells with bit-identical inputs, so the scheme
//      stays conservative without atomics or a separate face pass;
//   3. sums the sum-factorised divergence (three Q-point contractions) and
//      the six lifted gaps per solution point and scales by 1/det(J);
//   4. combines with the RK stage in the reference operation order and
//      streams the result out with coalesced flat stores.
#include "fr_hex.hpp"

#include "euler.hpp"

namespace mcfd {

namespace {

template <int Q>
struct Shape {
  static constexpr int kPoints = Q * Q * Q;
  static constexpr int kFacePoints = Q * Q;
  static constexpr int kCells = 256 / kPoints;  // cells per workgroup
  static_assert(kCells >= 1, "at most 256 solution points per cell");
};

__device__ __forceinline__ int LineNode(int q, int f, int a, int b, int m) {
  switch (f) {
  case 0: case 5: return (a * q + b) * q + m;
  case 1: case 3: return (a * q + m) * q + b;
  default: return (m * q + a) * q + b;
  }
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

template <int Q, int kRiemann>
__global__ __launch_bounds__(256) void FrHexStageKernel(FrHexArgs a, LineTables lt, BcTable bc) {
  using S = Shape<Q>;
  constexpr int NP = S::kPoints, NFP = S::kFacePoints, C = S::kCells;
  __shared__ double s_u[C * NP * 5];
  __shared__ double s_f[C * NP * 15];
  __shared__ double s_gap[C * 6 * NFP * 5];

  const int tid = threadIdx.x;
  const int64_t c0 = LogicalBlock(a.xcd_remap) * C;
  if (c0 >= a.n_cells) return;
  const int nc = (int)min<int64_t>(C, a.n_cells - c0);
  const int64_t base = c0 * NP * 5;

  // 0. flat load of the cells' states
  for (int e = tid; e < nc * NP * 5; e += 256) s_u[e] = a.us[base + e];
  __syncthreads();

  // 1. local flux F~ = F J* at every solution point
  for (int w = tid; w < nc * NP; w += 256) {
    const int cell = w / NP;
    const double *g = a.geo + (int64_t)a.geo_id[c0 + cell] * kGeoStride;
    State u;
#pragma unroll
    for (int k = 0; k < 5; ++k) u.v[k] = s_u[w * 5 + k];
    double f[5][3];
    FluxMatrix(u, f);
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        s_f[w * 15 + k * 3 + d] = f[k][0] * g[d] + f[k][1] * g[3 + d] + f[k][2] * g[6 + d];
  }
  __syncthreads();

  // 2. face points: traces, upwind flux, flux gap
  for (int w = tid; w < nc * 6 * NFP; w += 256) {
    const int cell = w / (6 * NFP);
    const int rem = w - cell * 6 * NFP;
    const int f = rem / NFP, fp = rem - f * NFP;
    const int ia = fp / Q, ib = fp - ia * Q;
    const int64_t cid = c0 + cell;
    const int side = (f == 2 || f == 3 || f == 5) ? 1 : 0;
    const int dir = (f == 2 || f == 4) ? 0 : ((f == 1 || f == 3) ? 1 : 2);
    const double *lw = lt.lend[side];
    State own;
    double fn[5];
    {
      double vals[Q], flux[Q];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int m = 0; m < Q; ++m) {
          int node = cell * NP + LineNode(Q, f, ia, ib, m);
          vals[m] = s_u[node * 5 + k];
          flux[m] = s_f[node * 15 + k * 3 + dir];
        }
        own.v[k] = Extrapolate<Q>(lw, vals);
        fn[k] = Extrapolate<Q>(lw, flux);
      }
    }
    const int32_t nb = a.nbr[cid * 6 + f];
    const int32_t info = a.finfo[cid * 6 + f];
    const double *fr = a.frames + (int64_t)(info >> 7) * kFrameStride;
    FaceFrame frame;
#pragma unroll
    for (int d = 0; d < 3; ++d) { frame.n[d] = fr[d]; frame.t[d] = fr[3 + d]; }
    frame.b[0] = frame.n[1] * frame.t[2] - frame.n[2] * frame.t[1];
    frame.b[1] = frame.n[2] * frame.t[0] - frame.n[0] * frame.t[2];
    frame.b[2] = frame.n[0] * frame.t[1] - frame.n[1] * frame.t[0];
    const double scale = a.geo[(int64_t)a.geo_id[cid] * kGeoStride + 10 + f];
    State fu;
    double sgn = 1.0;
    if (nb >= 0) {
      const int nf = info & 7, orient = (info >> 3) & 7;
      const bool holder = (info >> 6) & 1;
      int s = (orient & 1) ? ib : ia, t = (orient & 1) ? ia : ib;
      const int na = (orient & 2) ? Q - 1 - s : s;
      const int nbb = (orient & 4) ? Q - 1 - t : t;
      const int nside = (nf == 2 || nf == 3 || nf == 5) ? 1 : 0;
      const double *nw = lt.lend[nside];
      const double *src = a.us + (int64_t)nb * NP * 5;
      State other;
      double vals[Q];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int m = 0; m < Q; ++m) vals[m] = src[LineNode(Q, nf, na, nbb, m) * 5 + k];
        other.v[k] = Extrapolate<Q>(nw, vals);
      }
      fu = UpwindFlux<kRiemann>(frame, holder ? own : other, holder ? other : own);
      sgn = holder ? 1.0 : -1.0;
    } else {
      const int tag = -nb - 1;
      const int kind = bc.kind[tag];
      State given;
#pragma unroll
      for (int k = 0; k < 5; ++k) given.v[k] = bc.state[tag][k];
      if (kind == 1) {
        fu = WallFlux(frame, own);
      } else if (kind == 2) {
        fu = StateFlux(frame, own);
      } else if (kind == 3) {
        fu = StateFlux(frame, given);
      } else {
        fu = UpwindFlux<kRiemann>(frame, own, given);
      }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) s_gap[w * 5 + k] = fu.v[k] * (sgn * scale) - fn[k];
  }
  __syncthreads();

  // 3. divergence + lifted gaps + 1/det(J)
  double r[5];
  const bool active = tid < nc * NP;
  if (active) {
    const int cell = tid / NP, node = tid - cell * NP;
    const int i = node / (Q * Q), j = (node / Q) % Q, k = node % Q;
    const double *sf = s_f + cell * NP * 15;
    const double *sg = s_gap + cell * 6 * NFP * 5;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      double div = 0;
#pragma unroll
      for (int m = 0; m < Q; ++m) {
        div += lt.d[i][m] * sf[((m * Q + j) * Q + k) * 15 + v * 3 + 0];
        div += lt.d[j][m] * sf[((i * Q + m) * Q + k) * 15 + v * 3 + 1];
        div += lt.d[k][m] * sf[((i * Q + j) * Q + m) * 15 + v * 3 + 2];
      }
      double corr = sg[((0 * NFP) + i * Q + j) * 5 + v] * lt.gp[0][k]
                  + sg[((5 * NFP) + i * Q + j) * 5 + v] * lt.gp[1][k]
                  + sg[((1 * NFP) + i * Q + k) * 5 + v] * lt.gp[0][j]
                  + sg[((3 * NFP) + i * Q + k) * 5 + v] * lt.gp[1][j]
                  + sg[((4 * NFP) + j * Q + k) * 5 + v] * lt.gp[0][i]
                  + sg[((2 * NFP) + j * Q + k) * 5 + v] * lt.gp[1][i];
      r[v] = -div - corr;
    }
    const double inv_det = a.geo[(int64_t)a.geo_id[c0 + cell] * kGeoStride + 9];
#pragma unroll
    for (int v = 0; v < 5; ++v) r[v] *= inv_det;
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int v = 0; v < 5; ++v) s_f[tid * 5 + v] = r[v];
  }
  __syncthreads();

  // 4. RK stage in the reference order (temporal/rk.hpp:31-49, ode.hpp:121-128)
  const double dt = a.dt;
  for (int e = tid; e < nc * NP * 5; e += 256) {
    const double res = s_f[e], us = s_u[e];
    double out;
    if (a.stage == 0) {
      out = res;
    } else {
      const double next = __dadd_rn(us, __dmul_rn(res, dt));  // residual *= dt; u += residual
      if (a.stage == 1) {
        out = next;
      } else if (a.stage == 2) {
        out = __dadd_rn(__dmul_rn(a.u0[base + e], 3.0), next) / 4.0;
      } else {
        out = __dadd_rn(__dmul_rn(next, 2.0), a.u0[base + e]) / 3.0;
      }
    }
    a.out[base + e] = out;
  }
}

template <int Q>
__global__ __launch_bounds__(256) void FrHexTimeStepKernel(const double *u, const int32_t *geo_id,
    const double *geo, int64_t n_cells, unsigned long long *dt_min_bits) {
  constexpr int NP = Q * Q * Q;
  __shared__ double s_min[256];
  double local = 1e300;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < n_cells; c += (int64_t)gridDim.x * 256) {
    double vmax = 1e-5;  // fem.hpp:419, avoids 0 in the divisor
    const double *uc = u + c * NP * 5;
    for (int p = 0; p < NP; ++p) {
      State s;
#pragma unroll
      for (int k = 0; k < 5; ++k) s.v[k] = uc[p * 5 + k];
      vmax = fmax(vmax, MaxSpeed(s));
    }
    local = fmin(local, geo[(int64_t)geo_id[c] * kGeoStride + 16] / vmax);
  }
  s_min[threadIdx.x] = local;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMin(dt_min_bits, (unsigned long long)__double_as_longlong(s_min[0]));
}

template <int Q, int R>
hipError_t Launch(const FrHexArgs &args, const LineTables &lt, const BcTable &bc, hipStream_t stream) {
  constexpr int C = Shape<Q>::kCells;
  int64_t blocks = (args.n_cells + C - 1) / C;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL((FrHexStageKernel<Q, R>), dim3((unsigned)blocks), dim3(256), 0, stream, args, lt, bc);
  return hipGetLastError();
}

template <int Q>
hipError_t LaunchQ(int riemann, const FrHexArgs &args, const LineTables &lt, const BcTable &bc,
                   hipStream_t stream) {
  switch (riemann) {
  case kExact: return Launch<Q, kExact>(args, lt, bc, stream);
  case kHllc: return Launch<Q, kHllc>(args, lt, bc, stream);
  case kRoe: return Launch<Q, kRoe>(args, lt, bc, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t LaunchFrHexStage(int q, int riemann, const FrHexArgs &args, const LineTables &lt,
                            const BcTable &bc, hipStream_t stream) {
  switch (q) {
  case 2: return LaunchQ<2>(riemann, args, lt, bc, stream);
  case 3: return LaunchQ<3>(riemann, args, lt, bc, stream);
  case 4: return LaunchQ<4>(riemann, args, lt, bc, stream);
  case 5: return LaunchQ<5>(riemann, args, lt, bc, stream);
  case 6: return LaunchQ<6>(riemann, args, lt, bc, stream);
  default: return hipErrorInvalidValue;
  }
}

hipError_t LaunchFrHexTimeStep(int q, const double *u, const int32_t *geo_id, const double *geo,
                               int64_t n_cells, unsigned long long *dt_min_bits, hipStream_t stream) {
  int64_t blocks = std::min<int64_t>((n_cells + 255) / 256, 4096);
  if (blocks < 1) blocks = 1;
  switch (q) {
#define MCFD_TS_CASE(QQ) \
  case QQ: hipLaunchKernelGGL((FrHexTimeStepKernel<QQ>), dim3((unsigned)blocks), dim3(256), 0, stream, \
                              u, geo_id, geo, n_cells, dt_min_bits); break;
  MCFD_TS_CASE(2) MCFD_TS_CASE(3) MCFD_TS_CASE(4) MCFD_TS_CASE(5) MCFD_TS_CASE(6)
#undef MCFD_TS_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mcfd
