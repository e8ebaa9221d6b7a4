// The reference's requirements on a convective Riemann solver
// (riemann/concept.hpp:12-44), restated for the host surface: the types a
// scheme reads from it and the flux methods it calls.  C++20.
#ifndef MINI_RIEMANN_CONCEPT_HPP_
#define MINI_RIEMANN_CONCEPT_HPP_

#include <concepts>
#include <type_traits>

namespace mini {
namespace riemann {

template <typename R>
concept HasConvectiveData = requires {
  typename R::Convection;
  requires std::integral<decltype(R::kComponents)>;
  requires std::integral<decltype(R::kDimensions)>;
  typename R::Scalar;
  requires std::floating_point<typename R::Scalar>;
  typename R::Vector;
  typename R::Conservative;
  typename R::Flux;
  typename R::FluxMatrix;
  typename R::Jacobian;
};

template <typename R, typename C>
concept HasConvectiveMethods = requires(R riemann, C const &value) {
  requires HasConvectiveData<R>;
  requires std::same_as<C, typename R::Conservative>;
  { R::GetFluxMatrix(value) } -> std::same_as<typename R::FluxMatrix>;
  { riemann.GetFluxUpwind(value, value) } -> std::derived_from<typename R::Flux>;
  { R::GetMaximumSpeed(value) } -> std::same_as<typename R::Scalar>;
};

template <typename R>
concept Convective = HasConvectiveData<R> && HasConvectiveMethods<R, typename R::Conservative>;

}  // namespace riemann
}  // namespace mini

#endif  // MINI_RIEMANN_CONCEPT_HPP_
