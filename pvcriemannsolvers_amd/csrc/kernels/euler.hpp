// Device-only Euler helpers of the stage kernels (FP64), on top of the
// portable physics of mini/riemann/euler_core.hpp: the reciprocal-refined
// primitive recovery and Roe flux of the FR kernels, and the face flux as the
// FR kernels evaluate it.
#ifndef MINICFD_KERNELS_EULER_HPP_
#define MINICFD_KERNELS_EULER_HPP_

#include <hip/hip_runtime.h>

#include "../../../include/mini/riemann/euler_core.hpp"

namespace mcfd {

// ------------------------------------------------- fast FP64 reciprocal
// v_rcp_f64 refined by two Newton steps: within 1 ulp of 1/x for normal x.
// Used by the throughput kernels in place of IEEE division (a ~10
// instruction sequence on CDNA); both cells of a face run the same code, so
// the face flux stays bit-identical on the two sides.
__device__ __forceinline__ double FastRcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __fma_rn(-x, r, 1.0);
  r = __fma_rn(r, e, r);
  e = __fma_rn(-x, r, 1.0);
  return __fma_rn(r, e, r);
}

// Conservative -> primitive with the reference's zeroing rule
// (types.hpp:195-202,289-301) via one reciprocal of rho.  The plain
// arithmetic is done first; the rule only selects among its results, so
// a wavefront where no lane needs the rule (rho > 0, E >= 0, p >= 0 on every
// lane) skips the selects and gets bitwise the same values as one that
// applies them.  *inv_rho receives FastRcp of the returned density.
__device__ __forceinline__ State FastPrimitive(State const &c, double *inv_rho) {
  const double rho = c.v[0];
  const double inv = FastRcp(rho);
  State p;
  p.v[0] = rho;
  p.v[1] = c.v[1] * inv;
  p.v[2] = c.v[2] * inv;
  p.v[3] = c.v[3] * inv;
  const double ke = 0.5 * (p.v[1] * p.v[1] + p.v[2] * p.v[2] + p.v[3] * p.v[3]);
  p.v[4] = (c.v[4] - rho * ke) * kGm1;
  *inv_rho = inv;
  if (__all(rho > 0 && c.v[4] >= 0 && p.v[4] >= 0)) return p;
  // rho < 0 or E < 0: all zero; rho == 0: momenta and E kept as they are;
  // otherwise p < 0: all zero
  const bool bad = c.v[0] < 0 || c.v[4] < 0;
  const bool vacuum = !bad && c.v[0] == 0;
  const bool zero = bad || (!vacuum && p.v[4] < 0);
#pragma unroll
  for (int k = 1; k < 5; ++k) p.v[k] = zero ? 0.0 : (vacuum ? c.v[k] : p.v[k]);
  p.v[0] = zero ? 0.0 : rho;  // a vacuum keeps its signed zero
  if (zero || vacuum) *inv_rho = FastRcp(p.v[0]);
  return p;
}
__device__ __forceinline__ State FastPrimitive(State const &c) {
  double inv;
  return FastPrimitive(c, &inv);
}

// Roe flux of normal-frame primitives (same waves as RoeFlux) with
// reciprocals instead of divisions; inv_l, inv_r = FastRcp of the densities.
__device__ __forceinline__ State FastRoeFlux(State const &l, State const &r, double inv_l, double inv_r) {
  constexpr double kGoGm1 = kGamma / kGm1;
  double kel = 0.5 * (l.v[1] * l.v[1] + l.v[2] * l.v[2] + l.v[3] * l.v[3]);
  double ker = 0.5 * (r.v[1] * r.v[1] + r.v[2] * r.v[2] + r.v[3] * r.v[3]);
  double el = l.v[4] * (1 / kGm1) + l.v[0] * kel;  // total energy per volume
  double er = r.v[4] * (1 / kGm1) + r.v[0] * ker;
  State flux;
  flux.v[0] = 0.5 * (l.v[0] * l.v[1] + r.v[0] * r.v[1]);
  flux.v[1] = 0.5 * ((l.v[0] * l.v[1]) * l.v[1] + l.v[4] + (r.v[0] * r.v[1]) * r.v[1] + r.v[4]);
  flux.v[2] = 0.5 * ((l.v[0] * l.v[2]) * l.v[1] + (r.v[0] * r.v[2]) * r.v[1]);
  flux.v[3] = 0.5 * ((l.v[0] * l.v[3]) * l.v[1] + (r.v[0] * r.v[3]) * r.v[1]);
  flux.v[4] = 0.5 * ((el + l.v[4]) * l.v[1] + (er + r.v[4]) * r.v[1]);
  double sl = sqrt(l.v[0]), sr = sqrt(r.v[0]);
  double inv = FastRcp(sl + sr);
  double h0l = l.v[4] * inv_l * kGoGm1 + kel;
  double h0r = r.v[4] * inv_r * kGoGm1 + ker;
  double rho = sl * sr;
  double u = (sl * l.v[1] + sr * r.v[1]) * inv;
  double v = (sl * l.v[2] + sr * r.v[2]) * inv;
  double w = (sl * l.v[3] + sr * r.v[3]) * inv;
  double h = (sl * h0l + sr * h0r) * inv;
  double ke = (u * u + v * v + w * w) * 0.5;
  double aa = kGm1 * (h - ke);
  double a = sqrt(aa), ua = u * a, inv_aa = FastRcp(aa);
  double dp = r.v[4] - l.v[4], du = r.v[1] - l.v[1];
  double s1 = 0.25 * (dp - rho * a * du) * inv_aa * fabs(u - a);
  double s2 = 0.5 * ((r.v[0] - l.v[0]) - dp * inv_aa) * fabs(u);
  double s3 = 0.5 * (rho * (r.v[2] - l.v[2])) * fabs(u);
  double s4 = 0.5 * (rho * (r.v[3] - l.v[3])) * fabs(u);
  double s5 = 0.25 * (dp + rho * a * du) * inv_aa * fabs(u + a);
  flux.v[0] -= s1 + s2 + s5;
  flux.v[1] -= s1 * (u - a) + s2 * u + s5 * (u + a);
  flux.v[2] -= (s1 + s2 + s5) * v + s3;
  flux.v[3] -= (s1 + s2 + s5) * w + s4;
  flux.v[4] -= s1 * (h - ua) + s2 * ke + s3 * v + s4 * w + s5 * (h + ua);
  return flux;
}

// The upwind flux as the FR stage kernels evaluate it: Roe on primitives
// recovered with one reciprocal of rho per side (FastPrimitive, FastRoeFlux),
// the other solvers as UpwindFlux.
template <int kRiemann>
__device__ __forceinline__ State FaceFlux(FaceFrame const &f, State const &cl, State const &cr, bool &star) {
  if constexpr (kRiemann == kRoe) {
    double inv_l, inv_r;
    State l = FastPrimitive(cl, &inv_l), r = FastPrimitive(cr, &inv_r);
    GlobalToNormal(f, &l);
    GlobalToNormal(f, &r);
    State flux = FastRoeFlux(l, r, inv_l, inv_r);
    NormalToGlobal(f, &flux);
    return flux;
  } else {
    return UpwindFlux<kRiemann>(f, cl, cr, star);
  }
}

}  // namespace mcfd

#endif  // MINICFD_KERNELS_EULER_HPP_
