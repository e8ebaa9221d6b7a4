// mini::limiter::weno selection for the C++ host surface.
//
// The reference composes a limiter with a scheme as
// spatial::WithLimiter<Scheme, Limiter> (spatial/with_limiter.hpp:25-66)
// where the limiter is weno::Lazy (limiter/weno.hpp:99) or weno::Eigen
// (weno.hpp:199) and limiter::Reconstruct flags troubled cells with
// weno::Eigen::IsNotSmooth (weno.hpp:230-258).  The reconstructions run in
// device kernels; these tags carry the C-ABI identifiers.  kAll reconstructs
// every cell (Lazy::IsNotSmooth is always true, the demo's configuration).
#ifndef MINI_LIMITER_WENO_HPP_
#define MINI_LIMITER_WENO_HPP_

#include <cstdint>

#include "minicfd.h"

namespace mini {
namespace limiter {
namespace weno {

struct Lazy {
  static constexpr int32_t kFlagged = MCFD_LIMITER_WENO_FLAGGED;
  static constexpr int32_t kAll = MCFD_LIMITER_WENO_ALL;
};
struct Eigen {
  static constexpr int32_t kFlagged = MCFD_LIMITER_WENO_EIGEN;
  static constexpr int32_t kAll = MCFD_LIMITER_WENO_EIGEN_ALL;
};

}  // namespace weno
}  // namespace limiter
}  // namespace mini

#endif  // MINI_LIMITER_WENO_HPP_
