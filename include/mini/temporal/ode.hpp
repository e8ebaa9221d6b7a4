// mini::temporal ODE contract and the explicit Euler solver, C++ host surface.
//
// Same classes, methods and operation order as the reference's
// temporal/ode.hpp:17-134: System (four pure virtuals), Constant
// (dU/dt = A U), Solver and Euler.  A spatial scheme of this library
// (spatial/mi355x.hpp) is a System, so the reference's solvers -- or these --
// drive it unchanged.
#ifndef MINI_TEMPORAL_ODE_HPP_
#define MINI_TEMPORAL_ODE_HPP_

#include <cassert>

#include "mini/algebra/column.hpp"

namespace mini {
namespace temporal {

// temporal/ode.hpp:17-48
template <typename Scalar>
class System {
 public:
  using Column = algebra::DynamicVector<Scalar>;

  virtual ~System() = default;
  virtual void SetTime(double t_curr) = 0;
  virtual void SetSolutionColumn(Column const &) = 0;
  virtual Column GetSolutionColumn() const = 0;
  virtual Column GetResidualColumn() const = 0;
};

// temporal/ode.hpp:55-85: dU/dt = A U.
template <typename Scalar>
class Constant : public System<Scalar> {
 public:
  using Matrix = algebra::DynamicMatrix<Scalar>;
  using Column = typename System<Scalar>::Column;

  explicit Constant(Matrix const &a) : a_(a) { assert(a.rows() == a.cols()); }
  void SetTime(double) final {}
  void SetSolutionColumn(Column const &u) final {
    assert(u.size() == a_.cols());
    u_ = u;
  }
  Column GetSolutionColumn() const final { return u_; }
  Column GetResidualColumn() const final { return a_ * u_; }

 private:
  Matrix a_;
  Column u_;
};

// temporal/ode.hpp:92-105
template <typename Scalar>
class Solver {
 public:
  using Column = typename System<Scalar>::Column;
  virtual ~Solver() = default;
  virtual void Update(System<Scalar> *system, double t_curr, double dt) = 0;
};

// temporal/ode.hpp:112-134: U(t + dt) = U(t) + R(U(t)) dt.
template <typename Scalar>
class Euler : public Solver<Scalar> {
 public:
  using Column = typename Solver<Scalar>::Column;

  static Column NextSolution(System<Scalar> *system, double t_curr, double dt) {
    system->SetTime(t_curr);
    auto u_next = system->GetSolutionColumn();
    auto residual = system->GetResidualColumn();
    residual *= dt;
    u_next += residual;
    return u_next;
  }

  void Update(System<Scalar> *system, double t_curr, double dt) final {
    auto u_next = NextSolution(system, t_curr, dt);
    system->SetSolutionColumn(u_next);
  }
};

}  // namespace temporal
}  // namespace mini

#endif  // MINI_TEMPORAL_ODE_HPP_
