// C++ host surface, CPU only: mini::temporal::RungeKutta<1|2|3> advances a
// Constant system (dU/dt = A U) bit for bit as the closed-form Shu-Osher
// combinations evaluated element by element (the check of the reference's
// test/temporal/rk.cpp:21-72, with seeded random A, U, t and dt).
#include <cstdio>
#include <random>
#include <vector>

#include "mini/temporal/ode.hpp"
#include "mini/temporal/rk.hpp"

namespace {

template <typename Scalar>
struct Case {
  using System = mini::temporal::Constant<Scalar>;
  using Column = typename System::Column;
  typename System::Matrix a = typename System::Matrix(10, 10);
  Column u = Column(10);
  double t = 0, dt = 0;

  explicit Case(unsigned seed) {
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> sym(-1, 1), pos(0, 1);
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) a(i, j) = (Scalar)sym(gen);
    for (int i = 0; i < 10; ++i) u[i] = (Scalar)sym(gen);
    t = pos(gen);
    dt = pos(gen);
  }
  // v + (A v) dt, with dt a Scalar as in Euler::NextSolution (residual *= dt)
  Column Step(Column const &v) const {
    Column r = a * v, out(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      Scalar rdt = r[i];
      rdt *= (Scalar)dt;
      out[i] = v[i] + rdt;
    }
    return out;
  }
};

template <int kOrders, typename Scalar>
bool Check(unsigned seed) {
  Case<Scalar> c(seed);
  typename Case<Scalar>::System system(c.a);
  system.SetSolutionColumn(c.u);
  mini::temporal::RungeKutta<kOrders, Scalar> solver;
  solver.Update(&system, c.t, c.dt);
  auto got = system.GetSolutionColumn();
  auto u1 = c.Step(c.u);
  typename Case<Scalar>::Column want(c.u.size());
  if (kOrders == 1) {
    want = u1;
  } else if (kOrders == 2) {
    auto v = c.Step(u1);
    for (size_t i = 0; i < want.size(); ++i) want[i] = (v[i] + c.u[i]) * Scalar(0.5);
  } else {
    auto v = c.Step(u1);
    typename Case<Scalar>::Column u2(c.u.size());
    for (size_t i = 0; i < u2.size(); ++i) {
      Scalar x = c.u[i];
      x *= 3;
      x += v[i];
      x /= 4;
      u2[i] = x;
    }
    auto w = c.Step(u2);
    for (size_t i = 0; i < want.size(); ++i) {
      Scalar x = w[i];
      x *= 2;
      x += c.u[i];
      x /= 3;
      want[i] = x;
    }
  }
  const bool ok = got == want;
  std::printf("%s RungeKutta<%d, %s> seed %u\n", ok ? "PASS" : "FAIL", kOrders,
              sizeof(Scalar) == 4 ? "float" : "double", seed);
  return ok;
}

}  // namespace

int main() {
  bool ok = true;
  for (unsigned seed : { 31415926u, 2718281u, 1u }) {
    ok &= Check<1, float>(seed) && Check<2, float>(seed) && Check<3, float>(seed);
    ok &= Check<1, double>(seed) && Check<2, double>(seed) && Check<3, double>(seed);
  }
  return ok ? 0 : 1;
}
