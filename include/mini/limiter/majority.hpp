// limiter::majority::Reconstruct(part) (limiter/majority.hpp:14-42) for a
// nodal (flux-reconstruction) mini::spatial::mi355x::FiniteElement: per
// component, every solution point takes the value that more than half of the
// cell's points hold, on the device.  Degrees below 2 are left unchanged; a
// cell component with no majority throws (the reference asserts), as does a
// modal (DG) element.
#ifndef MINI_LIMITER_MAJORITY_HPP_
#define MINI_LIMITER_MAJORITY_HPP_

#include "mini/spatial/mi355x.hpp"

namespace mini {
namespace limiter {
namespace majority {

inline void Reconstruct(spatial::mi355x::FiniteElement *fe_ptr) {
  if (fe_ptr->Degree() < 2) {
    return;
  }
  spatial::mi355x::Check(mcfd_startup_limiter(fe_ptr->handle(), MCFD_STARTUP_MAJORITY));
}

}  // namespace majority
}  // namespace limiter
}  // namespace mini

#endif  // MINI_LIMITER_MAJORITY_HPP_
