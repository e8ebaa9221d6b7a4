// Face fluxes at independent points, evaluated by the same device code as
// the stage kernels: the upwind flux of rotated::Euler
// (riemann/rotated/euler.hpp:153-159) and the boundary fluxes of an inviscid
// wall (:160-174), a supersonic outlet (:144-151) and the subsonic inlet and
// outlet (:181-280).  It lets the device
// solvers be checked on the reference's known answers pair by pair.
#include "face_flux.hpp"

#include <mutex>

#include "../host/common.hpp"
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

// ---- mcfd_face_fluxes_noalloc: the same evaluation staged through a fixed
// device buffer instead of per-call allocations.
namespace {

constexpr int64_t kStageMax = MCFD_NOALLOC_MAX_POINTS;
// [frames n x 9][left n x 5][right n x 5][flux n x 5] doubles, then one byte
// of exact-solver state per point; each part sized for kStageMax points
constexpr size_t kOffLeft = size_t(kStageMax) * 9 * sizeof(double);
constexpr size_t kOffRight = kOffLeft + size_t(kStageMax) * 5 * sizeof(double);
constexpr size_t kOffFlux = kOffRight + size_t(kStageMax) * 5 * sizeof(double);
constexpr size_t kOffStar = kOffFlux + size_t(kStageMax) * 5 * sizeof(double);
constexpr size_t kScratchBytes = (kOffStar + size_t(kStageMax) + 255) / 256 * 256;

// Staging memory of the entry point.  Every call on a device shares that
// device's copy (it lives until the process ends), so calls that overlap on
// one GPU race; nothing in it has to start at zero, as every call writes what
// it reads.  Host code addresses it only through the per-device pointer from
// hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device with the device current.
unsigned char *ScratchOfCurrentDevice() {
  int dev = 0;
  mcfd::Check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw mcfd::Failure{ MCFD_ENOTSUP, "device index beyond the staging table" };
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    mcfd::Check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" int mcfd_face_fluxes_noalloc(int32_t riemann, int32_t form, int32_t kind, int64_t n,
                                        const double *frames, const double *left, const double *right,
                                        uint8_t *exact_star, double *flux) {
  using mcfd::Check;
  if (n < 1 || !frames || !left || !flux ||
      (kind != MCFD_FLUX_INVISCID_WALL && kind != MCFD_FLUX_SUPERSONIC_OUTLET && !right))
    return mcfd::Fail(MCFD_EINVAL, "null argument or empty point set");
  if (n > kStageMax) return mcfd::Fail(MCFD_EINVAL, "mcfd_face_fluxes_noalloc: more points than MCFD_NOALLOC_MAX_POINTS");
  if (riemann < MCFD_RIEMANN_EXACT || riemann > MCFD_RIEMANN_ROE) return mcfd::Fail(MCFD_EINVAL, "unknown Riemann solver");
  if (form != MCFD_FLUX_FORM_FR && form != MCFD_FLUX_FORM_DG) return mcfd::Fail(MCFD_EINVAL, "unknown flux form");
  if (kind < MCFD_FLUX_UPWIND || kind > MCFD_FLUX_SUBSONIC_OUTLET) return mcfd::Fail(MCFD_EINVAL, "unknown flux kind");
  return mcfd::Guard([&] {
    unsigned char *base = ScratchOfCurrentDevice();
    double *d_frames = reinterpret_cast<double *>(base);
    double *d_left = reinterpret_cast<double *>(base + kOffLeft);
    double *d_right = reinterpret_cast<double *>(base + kOffRight);
    double *d_flux = reinterpret_cast<double *>(base + kOffFlux);
    uint8_t *d_star = base + kOffStar;
    hipStream_t stream = nullptr;  // the device's default stream: no stream is created
    const size_t n9 = size_t(n) * 9 * sizeof(double), n5 = size_t(n) * 5 * sizeof(double);
    Check(hipMemcpyAsync(d_frames, frames, n9, hipMemcpyHostToDevice, stream), "copy");
    Check(hipMemcpyAsync(d_left, left, n5, hipMemcpyHostToDevice, stream), "copy");
    // the wall and supersonic-outlet kinds never use `right`: any valid array will do
    if (right) Check(hipMemcpyAsync(d_right, right, n5, hipMemcpyHostToDevice, stream), "copy");
    if (exact_star) Check(hipMemcpyAsync(d_star, exact_star, size_t(n), hipMemcpyHostToDevice, stream), "copy");
    Check(mcfd::LaunchFaceFluxPoints(riemann, form, kind, n, d_frames, d_left, right ? d_right : d_left,
                                     exact_star ? d_star : nullptr, d_flux, stream),
          "face flux launch");
    Check(hipMemcpyAsync(flux, d_flux, n5, hipMemcpyDeviceToHost, stream), "copy");
    if (exact_star) Check(hipMemcpyAsync(exact_star, d_star, size_t(n), hipMemcpyDeviceToHost, stream), "copy");
    Check(hipStreamSynchronize(stream), "sync");
  });
}
