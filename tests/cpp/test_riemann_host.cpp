// Host-side mini::riemann::rotated::Euler<Exact|HartenLaxLeerContact|Roe>:
// the Convective concept (riemann/concept.hpp:12-44) checked at compile time,
// and every flux method evaluated on the points given on stdin.
//
// stdin:  riemann n, then per point 9 frame rows (n, t, b), left, right,
//         subsonic-inlet given, subsonic-outlet given, second left, second
//         right (5 values each, conservative except the givens)
// stdout: per point 8 fluxes of 5 values: upwind, wall, supersonic outlet of
//         left, supersonic inlet of right, subsonic inlet, subsonic outlet,
//         smart boundary, and the upwind flux of the second pair evaluated by
//         the same solver object (the exact solver keeps its star state)
#include <cstdio>
#include <iostream>
#include <vector>

#include "mini/riemann/concept.hpp"
#include "mini/riemann/euler.hpp"

namespace {

using mini::riemann::euler::Exact;
using mini::riemann::euler::HartenLaxLeerContact;
using mini::riemann::euler::IdealGas;
using mini::riemann::euler::Roe;
template <class U>
using Rotated = mini::riemann::rotated::Euler<U>;

static_assert(mini::riemann::Convective<Rotated<Exact<IdealGas, 3>>>);
static_assert(mini::riemann::Convective<Rotated<HartenLaxLeerContact<IdealGas, 3>>>);
static_assert(mini::riemann::Convective<Rotated<Roe<IdealGas, 3>>>);
static_assert(Rotated<Exact<IdealGas, 3>>::kId == MCFD_RIEMANN_EXACT);
static_assert(Rotated<Roe<IdealGas, 3>>::kComponents == 5 && Rotated<Roe<IdealGas, 3>>::kDimensions == 3);

using Value = mini::riemann::euler::Value;

Value Read5() {
  Value v;
  for (double &x : v) std::cin >> x;
  return v;
}

void Print(Value const &v) {
  for (double x : v) std::printf(" %.17g", x);
}

template <class U>
int Run(int64_t n) {
  using R = Rotated<U>;
  std::vector<typename R::Frame> frames(n);
  std::vector<R> solvers(n);
  for (int64_t i = 0; i < n; ++i) {
    for (auto &row : frames[i])
      for (double &x : row) std::cin >> x;
    solvers[i].Rotate(frames[i]);
    const Value l = Read5(), r = Read5(), inlet = Read5(), outlet = Read5(), l2 = Read5(), r2 = Read5();
    if (!std::cin) {
      std::fprintf(stderr, "short input at point %lld\n", (long long)i);
      return 2;
    }
    R const &s = solvers[i];
    Print(s.GetFluxUpwind(l, r));
    Print(s.GetFluxOnInviscidWall(l));
    Print(s.GetFluxOnSupersonicOutlet(l));
    Print(s.GetFluxOnSupersonicInlet(r));
    Print(s.GetFluxOnSubsonicInlet(l, inlet));
    Print(s.GetFluxOnSubsonicOutlet(l, outlet));
    Print(s.GetFluxOnSmartBoundary(l, r));
    Print(s.GetFluxUpwind(l2, r2));
    std::printf("\n");
    if (&s.normal() != &frames[i][0]) return 3;
  }
  return 0;
}

}  // namespace

int main() {
  int riemann = -1;
  int64_t n = 0;
  std::cin >> riemann >> n;
  switch (riemann) {
    case MCFD_RIEMANN_EXACT: return Run<Exact<IdealGas, 3>>(n);
    case MCFD_RIEMANN_HLLC: return Run<HartenLaxLeerContact<IdealGas, 3>>(n);
    case MCFD_RIEMANN_ROE: return Run<Roe<IdealGas, 3>>(n);
    default: std::fprintf(stderr, "unknown solver %d\n", riemann); return 2;
  }
}
