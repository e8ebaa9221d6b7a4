// limiter::Reconstruct(part, limiter) (limiter/reconstruct.hpp:21-47) for a
// mini::spatial::mi355x::FiniteElement: one pass of the element's limiter
// over its current solution, on the device.  A reference driver calls it
// after Approximate (demo/euler/sourceless.cpp:139-147).  As in the
// reference, a degree-0 scheme or a null limiter is left unchanged.
#ifndef MINI_LIMITER_RECONSTRUCT_HPP_
#define MINI_LIMITER_RECONSTRUCT_HPP_

#include "mini/limiter/weno.hpp"
#include "mini/spatial/mi355x.hpp"

namespace mini {
namespace limiter {

template <class Limiter>
void Reconstruct(spatial::mi355x::FiniteElement *fe_ptr, Limiter *limiter_ptr) {
  if (!(fe_ptr && limiter_ptr) || fe_ptr->Degree() == 0) {
    return;
  }
  spatial::mi355x::Check(mcfd_limiter_reconstruct(fe_ptr->handle()));
}

}  // namespace limiter
}  // namespace mini

#endif  // MINI_LIMITER_RECONSTRUCT_HPP_
