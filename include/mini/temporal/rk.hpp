// mini::temporal::RungeKutta, C++ host surface: the Shu-Osher SSP schemes of
// orders 1, 2 and 3 with the operation order of the reference's
// temporal/rk.hpp:10-57, so a System advanced through it gets bitwise the
// columns the reference computes through the same System.
#ifndef MINI_TEMPORAL_RK_HPP_
#define MINI_TEMPORAL_RK_HPP_

#include "mini/temporal/ode.hpp"

namespace mini {
namespace temporal {

template <int kOrders, typename Scalar>
struct RungeKutta : public Solver<Scalar> {
  static_assert(kOrders >= 1 && kOrders <= 3, "SSP-RK of order 1, 2 or 3");
  using Column = typename Solver<Scalar>::Column;

  void Update(System<Scalar> *system, double t_curr, double dt) final {
    if constexpr (kOrders == 1) {
      euler_.Update(system, t_curr, dt);
    } else if constexpr (kOrders == 2) {
      auto u_curr = system->GetSolutionColumn();
      euler_.Update(system, t_curr, dt);
      auto u_next = euler_.NextSolution(system, t_curr + dt, dt);
      u_next += u_curr;
      u_next *= 0.5;
      system->SetSolutionColumn(u_next);
    } else {
      auto u_curr = system->GetSolutionColumn();
      euler_.Update(system, t_curr, dt);  // U1 = U0 + dt R(U0)
      auto u_next = u_curr;
      u_next *= 3;
      u_next += euler_.NextSolution(system, t_curr + dt, dt);
      u_next /= 4;  // U2 = (3 U0 + (U1 + dt R(U1))) / 4
      system->SetSolutionColumn(u_next);
      u_next = euler_.NextSolution(system, t_curr + dt / 2, dt);
      u_next *= 2;
      u_next += u_curr;
      u_next /= 3;  // U3 = (2 (U2 + dt R(U2)) + U0) / 3
      system->SetSolutionColumn(u_next);
    }
  }

 private:
  Euler<Scalar> euler_;
};

}  // namespace temporal
}  // namespace mini

#endif  // MINI_TEMPORAL_RK_HPP_
