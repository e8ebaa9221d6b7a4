// mini::riemann surface of the C++ host side, under the reference's names.
//
// The unrotated solvers euler::{Exact, HartenLaxLeerContact, Roe} act on
// normal-frame primitive pairs, and rotated::Euler<Unrotated> wraps one of
// them in a face point's frame, as riemann/euler/{exact,hllc}.hpp and
// riemann/rotated/euler.hpp do; rotated::Euler satisfies the reference's
// mini::riemann::Convective concept (riemann/concept.hpp:12-44, restated in
// mini/riemann/concept.hpp).  All of them run the portable physics of
// mini/riemann/euler_core.hpp, the code the device kernels are compiled from,
// so a host caller gets the kernels' arithmetic.  The same types select the
// device solver of a mini::spatial::mi355x::FiniteElement through kId.
//
// Values are std::array<double, 5>: conservative (rho, rho u, rho v, rho w,
// E) and primitive (rho, u, v, w, p); the reference's Eigen-based vectors
// hold the same components in the same order (riemann/euler/types.hpp:17-99).
#ifndef MINI_RIEMANN_EULER_HPP_
#define MINI_RIEMANN_EULER_HPP_

#include <array>
#include <cmath>
#include <cstdint>

#include "mini/riemann/euler_core.hpp"
#include "minicfd.h"

namespace mini {
namespace riemann {
namespace euler {

using Value = std::array<double, 5>;
using Vector = std::array<double, 3>;
using FluxMatrix = std::array<std::array<double, 3>, 5>;
using Jacobian = std::array<std::array<double, 5>, 5>;

namespace detail {
inline mcfd::State ToState(Value const &v) { return mcfd::State{ { v[0], v[1], v[2], v[3], v[4] } }; }
inline Value ToValue(mcfd::State const &s) { return Value{ s.v[0], s.v[1], s.v[2], s.v[3], s.v[4] }; }
}  // namespace detail

// riemann/euler/types.hpp:186-334 with gamma = 1.4 (the configs' gas)
struct IdealGas {
  using Scalar = double;
  static constexpr double Gamma() { return mcfd::kGamma; }
  static constexpr double R() { return mcfd::kGasR; }
  // types.hpp:195-202,289-301, with the zeroing rule for negative rho or p
  static Value ConservativeToPrimitive(Value const &c) { return detail::ToValue(mcfd::ToPrimitive(detail::ToState(c))); }
  static Value PrimitiveToConservative(Value const &p) { return detail::ToValue(mcfd::ToConservative(detail::ToState(p))); }
  // types.hpp:313-321: normal-frame flux of a primitive state
  static Value PrimitiveToFlux(Value const &p) { return detail::ToValue(mcfd::NormalFlux(detail::ToState(p))); }
  static double GetSpeedOfSound(Value const &p) { return mcfd::SoundSpeed(p[0], p[4]); }
};

// Common part of the unrotated solvers: types and the physical flux.
template <class G, int kId_>
struct Unrotated {
  static_assert(kId_ >= MCFD_RIEMANN_EXACT && kId_ <= MCFD_RIEMANN_ROE, "a solver of the library");
  using Gas = G;
  using Scalar = double;
  using Vector = euler::Vector;
  using Primitive = Value;
  using Conservative = Value;
  using Flux = Value;
  static constexpr int kComponents = 5;
  static constexpr int kDimensions = 3;
  static constexpr int32_t kId = kId_;  // the device solver (MCFD_RIEMANN_*)
  // GetFlux of exact.hpp:31-33 / hllc.hpp:25-27
  static Flux GetFlux(Primitive const &p) { return Gas::PrimitiveToFlux(p); }
  // the state of the solver object that persists between calls: the exact
  // solver keeps the sign of its last star velocity (exact.hpp:29)
  mutable bool star_positive = false;
  // GetFluxUpwind of exact.hpp:45-47 / hllc.hpp:30-69 on a normal-frame pair
  Flux GetFluxUpwind(Primitive const &l, Primitive const &r) const {
    const mcfd::State a = detail::ToState(l), b = detail::ToState(r);
    mcfd::State f;
    if constexpr (kId_ == MCFD_RIEMANN_EXACT) {
      f = mcfd::ExactFlux(a, b, star_positive);
    } else if constexpr (kId_ == MCFD_RIEMANN_HLLC) {
      f = mcfd::HllcFlux(a, b);
    } else {
      f = mcfd::RoeFlux(a, b);
    }
    return detail::ToValue(f);
  }
};

// riemann/euler/exact.hpp:15-321 (FindRoot stops at f >= -1e-8)
template <class Gas = IdealGas, int kDimensions = 3>
struct Exact : Unrotated<Gas, MCFD_RIEMANN_EXACT> {
  static_assert(kDimensions == 3, "the library implements 3-D states");
};
// riemann/euler/hllc.hpp:17-110
template <class Gas = IdealGas, int kDimensions = 3>
struct HartenLaxLeerContact : Unrotated<Gas, MCFD_RIEMANN_HLLC> {
  static_assert(kDimensions == 3, "the library implements 3-D states");
};
// python/riemann.py:405-451 (no entropy fix) with two shear waves in 3-D
template <class Gas = IdealGas, int kDimensions = 3>
struct Roe : Unrotated<Gas, MCFD_RIEMANN_ROE> {
  static_assert(kDimensions == 3, "the library implements 3-D states");
};

}  // namespace euler

namespace rotated {

// riemann/rotated/euler.hpp:24-356: an unrotated solver in a face point's
// orthonormal frame, rows (normal, tangent, bitangent).  One object per face
// point, as the reference's FiniteElement keeps them (fem.hpp:95-115).
template <class UnrotatedEuler>
class Euler {
 public:
  using Convection = Euler;
  static constexpr int kComponents = UnrotatedEuler::kComponents;
  static constexpr int kDimensions = UnrotatedEuler::kDimensions;
  static constexpr int32_t kId = UnrotatedEuler::kId;
  using Gas = typename UnrotatedEuler::Gas;
  using Scalar = typename UnrotatedEuler::Scalar;
  using Vector = typename UnrotatedEuler::Vector;
  using Frame = std::array<Vector, kDimensions>;
  using Flux = typename UnrotatedEuler::Flux;
  using FluxMatrix = euler::FluxMatrix;
  using Conservative = typename UnrotatedEuler::Conservative;
  using Primitive = typename UnrotatedEuler::Primitive;
  using Value = euler::Value;
  using Jacobian = euler::Jacobian;

  void Rotate(Frame const &frame) { frame_ = &frame; }
  Vector const &normal() const { return (*frame_)[0]; }

  // rotated/euler.hpp:117-120 (GetFluxMatrix of types.hpp:322-333)
  static FluxMatrix GetFluxMatrix(Conservative const &c) {
    double f[5][3];
    mcfd::FluxMatrix(euler::detail::ToState(c), f);
    FluxMatrix m;
    for (int k = 0; k < 5; ++k)
      for (int d = 0; d < 3; ++d) m[k][d] = f[k][d];
    return m;
  }
  // rotated/euler.hpp:340-345
  static Scalar GetMaximumSpeed(Conservative const &c) { return mcfd::MaxSpeed(euler::detail::ToState(c)); }
  // rotated/euler.hpp:153-159
  Flux GetFluxUpwind(Conservative const &left, Conservative const &right) const {
    mcfd::State l = mcfd::ToPrimitive(euler::detail::ToState(left));
    mcfd::State r = mcfd::ToPrimitive(euler::detail::ToState(right));
    const mcfd::FaceFrame f = FaceFrame();
    mcfd::GlobalToNormal(f, &l);
    mcfd::GlobalToNormal(f, &r);
    mcfd::State flux = euler::detail::ToState(unrotated_.GetFluxUpwind(euler::detail::ToValue(l),
                                                                        euler::detail::ToValue(r)));
    mcfd::NormalToGlobal(f, &flux);
    return euler::detail::ToValue(flux);
  }
  // rotated/euler.hpp:160-174, the default branch: (0, p n, 0)
  Flux GetFluxOnInviscidWall(Conservative const &interior) const {
    return euler::detail::ToValue(mcfd::WallFlux(FaceFrame(), euler::detail::ToState(interior)));
  }
  // rotated/euler.hpp:144-151
  Flux GetFluxOnSupersonicInlet(Conservative const &exterior) const {
    return euler::detail::ToValue(mcfd::StateFlux(FaceFrame(), euler::detail::ToState(exterior)));
  }
  Flux GetFluxOnSupersonicOutlet(Conservative const &interior) const {
    return euler::detail::ToValue(mcfd::StateFlux(FaceFrame(), euler::detail::ToState(interior)));
  }
  // rotated/euler.hpp:181-246: given = (p_total, u_cos, v_cos, w_cos, T_total)
  Flux GetFluxOnSubsonicInlet(Conservative const &interior, Value const &given) const {
    return euler::detail::ToValue(
        mcfd::SubsonicInletFlux(FaceFrame(), euler::detail::ToState(interior), euler::detail::ToState(given)));
  }
  // rotated/euler.hpp:263-280: given[4] = back pressure
  Flux GetFluxOnSubsonicOutlet(Conservative const &interior, Value const &given) const {
    return euler::detail::ToValue(
        mcfd::SubsonicOutletFlux(FaceFrame(), euler::detail::ToState(interior), euler::detail::ToState(given)));
  }
  // rotated/euler.hpp:296-299
  Flux GetFluxOnSmartBoundary(Conservative const &interior, Conservative const &exterior) const {
    return GetFluxUpwind(interior, exterior);
  }

 private:
  mcfd::FaceFrame FaceFrame() const {
    mcfd::FaceFrame f;
    for (int d = 0; d < 3; ++d) {
      f.n[d] = (*frame_)[0][d];
      f.t[d] = (*frame_)[1][d];
      f.b[d] = (*frame_)[2][d];
    }
    return f;
  }

  UnrotatedEuler unrotated_;
  Frame const *frame_ = nullptr;
};

}  // namespace rotated
}  // namespace riemann
}  // namespace mini

#endif  // MINI_RIEMANN_EULER_HPP_
