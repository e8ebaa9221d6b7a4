// limiter::average::Reconstruct(part) (limiter/average.hpp:11-28) for a
// nodal (flux-reconstruction) mini::spatial::mi355x::FiniteElement: every
// solution point of a cell takes the cell average, on the device.  Degree 0
// is left unchanged; a modal (DG) element throws, since the reference
// requires nodal polynomials (average.hpp:12).
#ifndef MINI_LIMITER_AVERAGE_HPP_
#define MINI_LIMITER_AVERAGE_HPP_

#include "mini/spatial/mi355x.hpp"

namespace mini {
namespace limiter {
namespace average {

inline void Reconstruct(spatial::mi355x::FiniteElement *fe_ptr) {
  if (fe_ptr->Degree() == 0) {
    return;
  }
  spatial::mi355x::Check(mcfd_startup_limiter(fe_ptr->handle(), MCFD_STARTUP_AVERAGE));
}

}  // namespace average
}  // namespace limiter
}  // namespace mini

#endif  // MINI_LIMITER_AVERAGE_HPP_
