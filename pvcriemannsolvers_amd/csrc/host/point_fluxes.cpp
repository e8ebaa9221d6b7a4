// C ABI of the point-wise face fluxes: the device Riemann solvers and
// boundary fluxes of the stage kernels, evaluated on caller-given state pairs.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../../include/minicfd.h"
#include "../kernels/face_flux.hpp"
#include "common.hpp"

using mcfd::Check;
using mcfd::Fail;
using mcfd::Guard;

extern "C" {

int mcfd_face_fluxes(int32_t riemann, int32_t form, int32_t kind, int64_t n, const double *frames,
                     const double *left, const double *right, uint8_t *exact_star, double *flux) {
  if (n < 1 || !frames || !left || !flux || (kind != MCFD_FLUX_INVISCID_WALL && kind != MCFD_FLUX_SUPERSONIC_OUTLET && !right))
    return Fail(MCFD_EINVAL, "null argument or empty point set");
  if (riemann < MCFD_RIEMANN_EXACT || riemann > MCFD_RIEMANN_ROE) return Fail(MCFD_EINVAL, "unknown Riemann solver");
  if (form != MCFD_FLUX_FORM_FR && form != MCFD_FLUX_FORM_DG) return Fail(MCFD_EINVAL, "unknown flux form");
  if (kind < MCFD_FLUX_UPWIND || kind > MCFD_FLUX_SUBSONIC_OUTLET) return Fail(MCFD_EINVAL, "unknown flux kind");
  return Guard([&] {
    hipStream_t stream;
    Check(hipStreamCreate(&stream), "hipStreamCreate");
    struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{ stream };
    mcfd::DeviceArray<double> d_frames, d_left, d_right, d_flux;
    mcfd::DeviceArray<uint8_t> d_star;
    d_frames.Allocate(size_t(n) * 9);
    d_left.Allocate(size_t(n) * 5);
    d_right.Allocate(size_t(n) * 5);
    d_flux.Allocate(size_t(n) * 5);
    Check(hipMemcpyAsync(d_frames.ptr, frames, size_t(n) * 9 * sizeof(double), hipMemcpyHostToDevice, stream), "copy");
    Check(hipMemcpyAsync(d_left.ptr, left, size_t(n) * 5 * sizeof(double), hipMemcpyHostToDevice, stream), "copy");
    if (right)
      Check(hipMemcpyAsync(d_right.ptr, right, size_t(n) * 5 * sizeof(double), hipMemcpyHostToDevice, stream), "copy");
    else
      Check(hipMemsetAsync(d_right.ptr, 0, size_t(n) * 5 * sizeof(double), stream), "memset");
    if (exact_star) {
      d_star.Allocate(size_t(n));
      Check(hipMemcpyAsync(d_star.ptr, exact_star, size_t(n), hipMemcpyHostToDevice, stream), "copy");
    }
    Check(mcfd::LaunchFaceFluxPoints(riemann, form, kind, n, d_frames.ptr, d_left.ptr, d_right.ptr, d_star.ptr,
                                     d_flux.ptr, stream),
          "face flux launch");
    Check(hipMemcpyAsync(flux, d_flux.ptr, size_t(n) * 5 * sizeof(double), hipMemcpyDeviceToHost, stream), "copy");
    if (exact_star)
      Check(hipMemcpyAsync(exact_star, d_star.ptr, size_t(n), hipMemcpyDeviceToHost, stream), "copy");
    Check(hipStreamSynchronize(stream), "sync");
  });
}

}  // extern "C"
