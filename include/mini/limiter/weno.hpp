// mini::limiter::weno::{Lazy, Eigen} of the C++ host surface.
//
// The reference composes a limiter with a scheme as
// spatial::WithLimiter<Scheme, Limiter>(limiter_ptr, ...)
// (spatial/with_limiter.hpp:19-66) and runs it through
// limiter::Reconstruct(part, limiter) (limiter/reconstruct.hpp:22-47).  Here
// the reconstructions run in device kernels of the FiniteElement they are
// attached to; a limiter object carries the reference constructor's
// parameters (linear weight w0, smoothness epsilon) and the C-ABI identifier
// of its troubled-cell rule:
//   Lazy   weno.hpp:99-195, IsNotSmooth always true (weno.hpp:122-124):
//          every cell is reconstructed (kAll); kFlagged restricts the lazy
//          reconstruction to the cells that Eigen's indicator flags;
//   Eigen  weno.hpp:199-348, characteristic-variable WENO of the cells that
//          IsNotSmooth (weno.hpp:230-258) flags (kFlagged); kAll reconstructs
//          every cell.
#ifndef MINI_LIMITER_WENO_HPP_
#define MINI_LIMITER_WENO_HPP_

#include <cstdint>

#include "minicfd.h"

namespace mini {
namespace limiter {
namespace weno {

class Lazy {
 public:
  using Scalar = double;
  static constexpr int32_t kFlagged = MCFD_LIMITER_WENO_FLAGGED;
  static constexpr int32_t kAll = MCFD_LIMITER_WENO_ALL;
  static constexpr int32_t kDefault = kAll;  // IsNotSmooth(cell) == true

  Lazy(Scalar w0, Scalar eps, bool verbose = false) : w0_(w0), eps_(eps), verbose_(verbose) {}
  Scalar w0() const { return w0_; }
  Scalar eps() const { return eps_; }
  bool verbose() const { return verbose_; }

 private:
  Scalar w0_, eps_;
  bool verbose_;
};

class Eigen {
 public:
  using Scalar = double;
  static constexpr int32_t kFlagged = MCFD_LIMITER_WENO_EIGEN;
  static constexpr int32_t kAll = MCFD_LIMITER_WENO_EIGEN_ALL;
  static constexpr int32_t kDefault = kFlagged;

  Eigen(Scalar w0, Scalar eps) : w0_(w0), eps_(eps) {}
  Scalar w0() const { return w0_; }
  Scalar eps() const { return eps_; }

 private:
  Scalar w0_, eps_;
};

}  // namespace weno
}  // namespace limiter
}  // namespace mini

#endif  // MINI_LIMITER_WENO_HPP_
