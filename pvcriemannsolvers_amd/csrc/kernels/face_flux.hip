// Face fluxes at independent points, evaluated by the same device code as
// the stage kernels: the upwind flux of rotated::Euler
// (riemann/rotated/euler.hpp:153-159) and the boundary fluxes of an inviscid
// wall (:160-174), a supersonic outlet (:144-151) and the subsonic inlet and
// outlet (:181-280).  It lets the device
// solvers be checked on the reference's known answers pair by pair.
#include "face_flux.hpp"

#include "euler.hpp"

namespace mcfd {

namespace {

template <int kRiemann, int kForm>
__global__ __launch_bounds__(64) void FaceFluxPointsKernel(int kind, int64_t n, const double *frames,
                                                            const double *left, const double *right,
                                                            uint8_t *star_state, double *flux) {
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  FaceFrame f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    f.n[d] = frames[i * 9 + d];
    f.t[d] = frames[i * 9 + 3 + d];
    f.b[d] = frames[i * 9 + 6 + d];
  }
  State l, r, out;
#pragma unroll
  for (int k = 0; k < 5; ++k) { l.v[k] = left[i * 5 + k]; r.v[k] = right[i * 5 + k]; }
  bool star = star_state ? star_state[i] != 0 : false;
  if (kind == kFluxInviscidWall) {
    out = WallFlux(f, l);
  } else if (kind == kFluxSupersonicOutlet) {
    out = StateFlux(f, l);
  } else if (kind == kFluxSubsonicInlet) {
    out = SubsonicInletFlux(f, l, r);
  } else if (kind == kFluxSubsonicOutlet) {
    out = SubsonicOutletFlux(f, l, r);
  } else if constexpr (kForm == 0) {
    out = FaceFlux<kRiemann>(f, l, r, star);
  } else {
    out = UpwindFlux<kRiemann>(f, l, r, star);
  }
  if (star_state) star_state[i] = star ? 1 : 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) flux[i * 5 + k] = out.v[k];
}

template <int kForm>
hipError_t LaunchForm(int riemann, int kind, int64_t n, const double *frames, const double *left,
                      const double *right, uint8_t *star_state, double *flux, hipStream_t stream) {
  const dim3 grid((unsigned)((n + 63) / 64)), block(64);
  switch (riemann) {
  case kExact:
    hipLaunchKernelGGL((FaceFluxPointsKernel<kExact, kForm>), grid, block, 0, stream, kind, n, frames, left, right,
                       star_state, flux);
    break;
  case kHllc:
    hipLaunchKernelGGL((FaceFluxPointsKernel<kHllc, kForm>), grid, block, 0, stream, kind, n, frames, left, right,
                       star_state, flux);
    break;
  case kRoe:
    hipLaunchKernelGGL((FaceFluxPointsKernel<kRoe, kForm>), grid, block, 0, stream, kind, n, frames, left, right,
                       star_state, flux);
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t LaunchFaceFluxPoints(int riemann, int form, int kind, int64_t n, const double *frames,
                                const double *left, const double *right, uint8_t *star_state, double *flux,
                                hipStream_t stream) {
  if (n < 1 || n > (int64_t)0x7fffffff * 64) return hipErrorInvalidValue;
  if (form == 0) return LaunchForm<0>(riemann, kind, n, frames, left, right, star_state, flux, stream);
  if (form == 1) return LaunchForm<1>(riemann, kind, n, frames, left, right, star_state, flux, stream);
  return hipErrorInvalidValue;
}

}  // namespace mcfd
