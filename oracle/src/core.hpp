// Oracle: CPU restatement of miniCFD's reference arithmetic.
//
// THIS DIRECTORY IS TEST INFRASTRUCTURE.  Only tests/, the smoke check and the
// CPU-baseline leg of bench.py may load it, and only as the checker.  The
// product path (pvcriemannsolvers_amd) never links or calls anything here.
//
// The reference C++ (include/mini/**) cannot be compiled in this image: every
// header includes "Eigen/Eigen" (include/mini/algebra/eigen.hpp:9), an
// un-vendored git submodule (ThirdParty/eigen, branch master, no pinned
// version).  This file restates the arithmetic loop by loop in plain C++ with
// fixed-size arrays, citing the reference file:line each piece follows.
//
// Parity pins (checked in tests/test_oracle_*.py):
//   * quadrature: monomial exactness (test/integrator/{legendre,lobatto}.cpp);
//   * Lagrange basis: Kronecker delta / partition of unity / derivative
//     identities (test/basis/lagrange.cpp);
//   * Vincent g: end values, orthogonality, symmetry (test/basis/vincent.cpp);
//   * exact Riemann: Toro known answers (test/riemann/euler/exact.cpp);
//   * HLLC: consistency F(u,u)=F(u) and closeness to the exact states
//     (test/riemann/euler/hllc.cpp);
//   * Roe: consistency, exact resolution of an isolated Rankine-Hugoniot
//     shock and of a stationary contact, mirror symmetry.  No reference
//     golden vectors exist (python/riemann.py is 1-D only and was not run),
//     so Roe flux values are parity unpinned beyond these invariants.
#ifndef ORACLE_CORE_HPP_
#define ORACLE_CORE_HPP_

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace oracle {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;  // [row][col]
using Val5 = std::array<double, 5>;
using Flux53 = std::array<std::array<double, 3>, 5>;  // [component][xyz]

inline Vec3 sub(Vec3 a, Vec3 const &b) { for (int i = 0; i < 3; ++i) a[i] -= b[i]; return a; }
inline Vec3 add(Vec3 a, Vec3 const &b) { for (int i = 0; i < 3; ++i) a[i] += b[i]; return a; }
inline Vec3 scale(Vec3 a, double s) { for (auto &x : a) x *= s; return a; }
inline double dot(Vec3 const &a, Vec3 const &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm(Vec3 const &a) { return std::sqrt(dot(a, a)); }
inline Vec3 cross(Vec3 const &a, Vec3 const &b) {
  return { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
}
inline Vec3 normalized(Vec3 const &a) { return scale(a, 1.0 / norm(a)); }

inline double det3(Mat3 const &m) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
       - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
       + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
inline Mat3 inv3(Mat3 const &m) {
  double d = det3(m);
  Mat3 r;
  r[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / d;
  r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / d;
  r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / d;
  r[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / d;
  r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / d;
  r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / d;
  r[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / d;
  r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / d;
  r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / d;
  return r;
}

// ---------------------------------------------------------------- quadrature
// 1-D Gauss-Legendre rules, closed forms of integrator/legendre.hpp:36-108.
inline void LegendreRule(int q, double *x, double *w) {
  switch (q) {
  case 1: x[0] = 0; w[0] = 2; break;
  case 2: x[0] = -std::sqrt(1.0 / 3.0); x[1] = -x[0]; w[0] = w[1] = 1; break;
  case 3: x[0] = -std::sqrt(0.6); x[1] = 0; x[2] = -x[0];
          w[0] = w[2] = 5.0 / 9.0; w[1] = 8.0 / 9.0; break;
  case 4: {
    double a = std::sqrt((3 + 2 * std::sqrt(1.2)) / 7), b = std::sqrt((3 - 2 * std::sqrt(1.2)) / 7);
    x[0] = -a; x[1] = -b; x[2] = b; x[3] = a;
    w[0] = w[3] = (18 - std::sqrt(30.0)) / 36; w[1] = w[2] = (18 + std::sqrt(30.0)) / 36;
    break;
  }
  case 5: {
    double a = std::sqrt((5 + std::sqrt(40 / 7.0)) / 9), b = std::sqrt((5 - std::sqrt(40 / 7.0)) / 9);
    x[0] = -a; x[1] = -b; x[2] = 0; x[3] = b; x[4] = a;
    w[0] = w[4] = (322 - 13 * std::sqrt(70.0)) / 900;
    w[1] = w[3] = (322 + 13 * std::sqrt(70.0)) / 900; w[2] = 128.0 / 225.0;
    break;
  }
  case 6: {
    const double p[3] = { 0.932469514203152, 0.6612093864662645, 0.23861918608319693 };
    const double v[3] = { 0.17132449237917016, 0.36076157304813855, 0.4679139345726912 };
    for (int i = 0; i < 3; ++i) { x[i] = -p[i]; x[5 - i] = p[i]; w[i] = w[5 - i] = v[i]; }
    break;
  }
  default: throw std::invalid_argument("Legendre rule: 1 <= Q <= 6");
  }
}

// 1-D Gauss-Lobatto rules, closed forms of integrator/lobatto.hpp:37-80.
inline void LobattoRule(int q, double *x, double *w) {
  switch (q) {
  case 2: x[0] = -1; x[1] = 1; w[0] = w[1] = 1; break;
  case 3: x[0] = -1; x[1] = 0; x[2] = 1; w[0] = w[2] = 1.0 / 3; w[1] = 4.0 / 3; break;
  case 4: x[0] = -1; x[1] = -std::sqrt(1.0 / 5); x[2] = -x[1]; x[3] = 1;
          w[0] = w[3] = 1.0 / 6; w[1] = w[2] = 5.0 / 6; break;
  case 5: x[0] = -1; x[1] = -std::sqrt(21.0) / 7; x[2] = 0; x[3] = -x[1]; x[4] = 1;
          w[0] = w[4] = 0.1; w[1] = w[3] = 49.0 / 90; w[2] = 32.0 / 45; break;
  case 6: {
    double a = std::sqrt((7 + 2 * std::sqrt(7.0)) / 21), b = std::sqrt((7 - 2 * std::sqrt(7.0)) / 21);
    x[0] = -1; x[1] = -a; x[2] = -b; x[3] = b; x[4] = a; x[5] = 1;
    w[0] = w[5] = 1.0 / 15; w[1] = w[4] = (14.0 - std::sqrt(7.0)) / 30;
    w[2] = w[3] = (14.0 + std::sqrt(7.0)) / 30;
    break;
  }
  default: throw std::invalid_argument("Lobatto rule: 2 <= Q <= 6");
  }
}

// ------------------------------------------------------------ Lagrange basis
// 1-D Lagrange basis on nodes x[0..q), basis/lagrange.hpp:34-145.  The
// reference differentiates through the inverse Taylor->Lagrange matrix
// (lagrange.hpp:72-85); the product rule used here is the same polynomial.
struct Lagrange1D {
  int q;
  double x[8];
  double Value(int i, double s) const {
    double v = 1;
    for (int m = 0; m < q; ++m) if (m != i) v *= (s - x[m]) / (x[i] - x[m]);
    return v;
  }
  double Derivative(int i, double s) const {
    double sum = 0;
    for (int l = 0; l < q; ++l) {
      if (l == i) continue;
      double t = 1.0 / (x[i] - x[l]);
      for (int m = 0; m < q; ++m) if (m != i && m != l) t *= (s - x[m]) / (x[i] - x[m]);
      sum += t;
    }
    return sum;
  }
};

// ----------------------------------------------------------- Vincent (ESFR)
// basis/vincent.hpp:19-102: g_right = (P_k + c_prev P_{k-1} + c_next P_{k+1})/2.
struct Vincent {
  int degree;
  double c_prev, c_next;
  Vincent(int k, double c_n) : degree(k), c_prev(1 - c_n), c_next(c_n) {}
  double RightValue(double s) const {
    double v = std::legendre(degree, s);
    v += c_prev * std::legendre(degree - 1, s);
    v += c_next * std::legendre(degree + 1, s);
    return 0.5 * v;
  }
  double LeftValue(double s) const { return RightValue(-s); }
  // vincent.hpp:75-91, the P'_{k+1} = (k+1) P_k + x P'_k recurrence.
  double RightDerivative(double s) const {
    double d_prev = 0, d_curr = 0, d_next = 0;
    for (int k = 0; k <= degree; ++k) {
      if (k > 0) { d_prev = d_curr; d_curr = d_next; }
      d_next = (k + 1) * std::legendre(k, s) + s * d_curr;
    }
    return 0.5 * (d_curr + c_prev * d_prev + c_next * d_next);
  }
  double LeftDerivative(double s) const { return -RightDerivative(-s); }
};

// ------------------------------------------------------------------ ideal gas
// riemann/euler/types.hpp:186-334 with gamma = 1.4.
struct Gas {
  static constexpr double kGamma = 1.4;
  static constexpr double Gamma() { return kGamma; }
  static constexpr double OneOverGamma() { return 1 / kGamma; }
  static constexpr double GammaPlusOne() { return kGamma + 1; }
  static constexpr double GammaPlusOneOverTwo() { return GammaPlusOne() / 2; }
  static constexpr double GammaPlusOneOverFour() { return GammaPlusOne() / 4; }
  static constexpr double GammaMinusOne() { return kGamma - 1; }
  static constexpr double OneOverGammaMinusOne() { return 1 / GammaMinusOne(); }
  static constexpr double GammaMinusOneOverTwo() { return GammaMinusOne() / 2; }
  static constexpr double GammaMinusOneUnderTwo() { return 2 / GammaMinusOne(); }
  // types.hpp:252-254
  static double SpeedOfSound(double rho, double p) {
    return (rho > 0 && p > 0) ? std::sqrt(Gamma() * p / rho) : 0;
  }
};

// Primitive (rho, u, v, w, p) stored in a Val5.
// types.hpp:195-202 SetZeroIfNegative; :289-301 ConservativeToPrimitive.
inline Val5 ConservativeToPrimitive(Val5 const &cv) {
  Val5 pv = cv;
  auto zero_if_negative = [](Val5 *x) {
    if ((*x)[0] < 0) x->fill(0);
    if ((*x)[4] < 0) x->fill(0);
  };
  zero_if_negative(&pv);
  if (pv[0]) {
    double rho = pv[0];
    pv[1] /= rho; pv[2] /= rho; pv[3] /= rho;
    double ke = 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
    pv[4] -= rho * ke;
    pv[4] *= Gas::GammaMinusOne();
    zero_if_negative(&pv);
  }
  return pv;
}
// types.hpp:303-311
inline Val5 PrimitiveToConservative(Val5 const &pv) {
  Val5 cv = pv;
  cv[1] *= pv[0]; cv[2] *= pv[0]; cv[3] *= pv[0];
  cv[4] *= Gas::OneOverGammaMinusOne();
  cv[4] += pv[0] * (0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]));
  return cv;
}
// types.hpp:313-321: the flux in the x direction of a primitive state.
inline Val5 PrimitiveToFlux(Val5 const &pv) {
  Val5 f = PrimitiveToConservative(pv);
  for (auto &x : f) x *= pv[1];
  f[1] += pv[4];
  f[4] += pv[4] * pv[1];
  return f;
}
// types.hpp:322-333: the 5x3 physical flux matrix of a conservative state.
inline Flux53 GetFluxMatrix(Val5 const &cv) {
  Val5 pv = ConservativeToPrimitive(cv);
  double u = pv[1], v = pv[2], w = pv[3], p = pv[4];
  double ru = cv[1], rv = cv[2], rw = cv[3], rh0 = cv[4] + p;
  Flux53 m;
  m[0] = { ru, rv, rw };
  m[1] = { ru * u + p, ru * v, ru * w };
  m[2] = { rv * u, rv * v + p, rv * w };
  m[3] = { rw * u, rw * v, rw * w + p };
  m[4] = { rh0 * u, rh0 * v, rh0 * w };
  return m;
}
// riemann/rotated/euler.hpp:340-345
inline double GetMaximumSpeed(Val5 const &cv) {
  Val5 pv = ConservativeToPrimitive(cv);
  double c = Gas::SpeedOfSound(pv[0], pv[4]);
  double u = std::sqrt(pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
  return std::max(c + u, c - u);
}

// ------------------------------------------------------------ exact solver
// riemann/euler/exact.hpp:15-321 (Toro ch. 4).  Primitives are in the normal
// frame: [1] is the normal velocity, [2],[3] the passive tangential ones.
struct ExactSolver {
  double star_u = 0.0;  // exact.hpp:30; kept across calls like the reference

  struct SpeedChange {  // exact.hpp:153-200
    double rho, p, a, p_const;
    explicit SpeedChange(Val5 const &b)
        : rho(b[0]), p(b[4]), a(b[0] == 0 ? 0 : Gas::SpeedOfSound(b[0], b[4])),
          p_const(b[4] * Gas::GammaMinusOneOverTwo()) {}
    double P(double pa) const { return Gas::GammaPlusOneOverTwo() * pa + p_const; }
    double operator()(double pa) const {
      if (pa >= p) return (pa - p) / std::sqrt(rho * P(pa));
      constexpr double e = Gas::GammaMinusOneOverTwo() / Gas::Gamma();
      return (std::pow(pa / p, e) - 1) * (Gas::GammaMinusOneUnderTwo() * a);
    }
    double Prime(double pa) const {
      if (pa >= p) {
        double pp = P(pa);
        double v = Gas::GammaPlusOneOverFour() * (p - pa);
        return (1 + v / pp) / std::sqrt(rho * pp);
      }
      constexpr double e = -Gas::GammaPlusOneOverTwo() / Gas::Gamma();
      return std::pow(pa / p, e) / (rho * a);
    }
  };
  static Val5 InsideExpansion(double gri_1, double gri_2) {  // exact.hpp:255-266
    constexpr double r = Gas::GammaMinusOne() / Gas::GammaPlusOne();
    double a = r * gri_2, a2 = a * a;
    double rho = std::pow(a2 / gri_1 * Gas::OneOverGamma(), Gas::OneOverGammaMinusOne());
    return { rho, a, 0, 0, a2 * rho * Gas::OneOverGamma() };
  }
  template <int kField>
  static Val5 NearShock(Val5 const &before, Val5 after) {  // exact.hpp:202-234
    double su = before[1];
    double divisor = (after[1] - before[1]) * before[0];
    su += (divisor ? (after[4] - before[4]) / divisor : before[1]);
    bool after_shock = (kField == 1) ? (su < 0) : (su > 0);
    if (after_shock) {
      after[0] = before[0];
      after[0] *= (before[1] - su) / (after[1] - su);
      return after;
    }
    return before;
  }
  template <int kField>
  static Val5 NearExpansion(Val5 const &before, Val5 after) {  // exact.hpp:236-292
    auto pm = [](double x, double y) { return kField == 1 ? x + y : x - y; };
    double a_before = Gas::SpeedOfSound(before[0], before[4]);
    double gri_1 = before[4] / std::pow(before[0], Gas::Gamma());
    double gri_2 = pm(before[1], Gas::GammaMinusOneUnderTwo() * a_before);
    double a_after = pm(a_before, Gas::GammaMinusOneOverTwo() * (before[1] - after[1]));
    bool axis_after = kField == 1 ? (after[1] - a_after < 0) : (after[1] + a_after > 0);
    bool axis_before = kField == 1 ? (before[1] - a_before > 0) : (before[1] + a_before < 0);
    if (axis_after) {
      after[0] = Gas::Gamma() * after[4] / (a_after * a_after);
      return after;
    } else if (axis_before) {
      return before;
    }
    return InsideExpansion(gri_1, gri_2);
  }
  static Val5 NearVacuum(Val5 const &l, Val5 const &r) {  // exact.hpp:293-320
    double la = Gas::SpeedOfSound(l[0], l[4]);
    if (l[1] > la) return l;
    double ra = Gas::SpeedOfSound(r[0], r[4]);
    if (r[1] + ra < 0) return r;
    double gri_1 = l[4] / std::pow(l[0], Gas::Gamma());
    double gri_2 = l[1] + la * Gas::GammaMinusOneUnderTwo();
    if (gri_2 >= 0) return InsideExpansion(gri_1, gri_2);
    gri_1 = r[4] / std::pow(r[0], Gas::Gamma());
    gri_2 = r[1] - ra * Gas::GammaMinusOneUnderTwo();
    if (gri_2 < 0) return InsideExpansion(gri_1, gri_2);
    return Val5{ 0, 0, 0, 0, 0 };
  }
  // exact.hpp:62-138 (without the passive scalars)
  Val5 PrimitiveUpwindCore(Val5 const &l, Val5 const &r) {
    Val5 result{ 0, 0, 0, 0, 0 };
    if (l[0] > 0) {
      if (!(r[0] > 0)) {  // only right vacuum
        double al = Gas::SpeedOfSound(l[0], l[4]);
        if (l[1] >= al) return l;
        double us = l[1] + al * Gas::GammaMinusOneUnderTwo();
        if (us <= 0) return result;
        return InsideExpansion(l[4] / std::pow(l[0], Gas::Gamma()), us);
      }
    } else {
      if (r[0] > 0) {  // only left vacuum
        double ar = Gas::SpeedOfSound(r[0], r[4]);
        if (r[1] + ar <= 0) return r;
        double us = r[1] - ar * Gas::GammaMinusOneUnderTwo();
        if (us >= 0) return result;
        return InsideExpansion(r[4] / std::pow(r[0], Gas::Gamma()), us);
      }
      return result;  // both vacuum
    }
    SpeedChange cl(l), cr(r);
    double du = r[1] - l[1];
    auto f = [&](double p) { return cl(p) + cr(p) + du; };
    auto fp = [&](double p) { return cl.Prime(p) + cr.Prime(p); };
    if (f(0) < 0) {
      // FindRoot, exact.hpp:140-152
      double x = l[4];
      while (f(x) > 0) x *= 0.5;
      while (f(x) < -1e-8) x -= f(x) / fp(x);
      Val5 star{ 0, 0, 0, 0, x };
      star[1] = 0.5 * (r[1] + cr(x) + l[1] - cl(x));
      star_u = star[1];
      if (0 < star[1]) {
        result = (x >= l[4]) ? NearShock<1>(l, star) : NearExpansion<1>(l, star);
      } else {
        result = (x >= r[4]) ? NearShock<3>(r, star) : NearExpansion<3>(r, star);
      }
    } else {
      result = NearVacuum(l, r);
    }
    return result;
  }
  Val5 PrimitiveUpwind(Val5 const &l, Val5 const &r) {  // exact.hpp:36-43,49-60
    Val5 s = PrimitiveUpwindCore(l, r);
    s[2] = star_u > 0 ? l[2] : r[2];
    s[3] = star_u > 0 ? l[3] : r[3];
    return s;
  }
  Val5 FluxUpwind(Val5 const &l, Val5 const &r) { return PrimitiveToFlux(PrimitiveUpwind(l, r)); }
};

// ---------------------------------------------------------------------- HLLC
// riemann/euler/hllc.hpp:30-109 (Toro ch. 10, PVRS + q factors).
struct HllcSolver {
  static double Q(double p_est, double pk) {
    if (p_est <= pk) return 1.0;
    return std::sqrt(1 + Gas::GammaPlusOneOverTwo() * (p_est / pk - 1) / Gas::Gamma());
  }
  static Val5 StarFlux(Val5 const &s, double wk, double ws) {
    Val5 flux = PrimitiveToFlux(s);
    Val5 uk = PrimitiveToConservative(s);
    double energy = uk[4];
    double temp = s[0] * (wk - s[1]) / (wk - ws);
    Val5 us;
    us[0] = temp;
    us[1] = temp * ws; us[2] = temp * s[2]; us[3] = temp * s[3];
    us[4] = energy / s[0] + (ws - s[1]) * (ws + s[4] / (s[0] * (wk - s[1])));
    us[4] *= temp;
    for (int i = 0; i < 5; ++i) { us[i] -= uk[i]; us[i] *= wk; flux[i] += us[i]; }
    return flux;
  }
  Val5 FluxUpwind(Val5 const &l, Val5 const &r) {
    double rho_avg = (l[0] + r[0]) / 2;
    double al = Gas::SpeedOfSound(l[0], l[4]), ar = Gas::SpeedOfSound(r[0], r[4]);
    double a_avg = (al + ar) / 2;
    double p_pvrs = (l[4] + r[4] - (r[1] - l[1]) * rho_avg * a_avg) / 2;
    double p_est = std::max(0.0, p_pvrs);
    double wl = l[1] - al * Q(p_est, l[4]);
    double wr = r[1] + ar * Q(p_est, r[4]);
    double ws = (r[4] - l[4] + l[0] * l[1] * (wl - l[1]) - r[0] * r[1] * (wr - r[1]))
              / (l[0] * (wl - l[1]) - r[0] * (wr - r[1]));
    if (0.0 <= wl) return PrimitiveToFlux(l);
    if (wr <= 0.0) return PrimitiveToFlux(r);
    if (0.0 <= ws) return StarFlux(l, wl, ws);
    return StarFlux(r, wr, ws);
  }
};

// ----------------------------------------------------------------------- Roe
// python/riemann.py:414-451 (1-D Roe without entropy fix), extended to 3-D
// in the rotated frame with two shear waves carrying the passive tangential
// velocities (Toro, "Riemann Solvers and Numerical Methods", sec. 11.2).
struct RoeSolver {
  Val5 FluxUpwind(Val5 const &l, Val5 const &r) {
    Val5 fl = PrimitiveToFlux(l), fr = PrimitiveToFlux(r);
    Val5 flux;
    for (int i = 0; i < 5; ++i) flux[i] = (fl[i] + fr[i]) / 2;
    auto h0 = [](Val5 const &s) {
      double ke = 0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]);
      return s[4] / s[0] * (Gas::Gamma() / Gas::GammaMinusOne()) + ke;
    };
    double sl = std::sqrt(l[0]), sr = std::sqrt(r[0]), ss = sl + sr;
    double rho = sl * sr;
    double u = (sl * l[1] + sr * r[1]) / ss;
    double v = (sl * l[2] + sr * r[2]) / ss;
    double w = (sl * l[3] + sr * r[3]) / ss;
    double h = (sl * h0(l) + sr * h0(r)) / ss;
    double ke = (u * u + v * v + w * w) / 2;
    double aa = Gas::GammaMinusOne() * (h - ke);
    double a = std::sqrt(aa), ua = u * a;
    double dp = r[4] - l[4], du = r[1] - l[1];
    // acoustic wave u - a
    double alpha = (dp - rho * a * du) / (2 * aa), s = 0.5 * alpha * std::abs(u - a);
    flux[0] -= s; flux[1] -= s * (u - a); flux[2] -= s * v; flux[3] -= s * w; flux[4] -= s * (h - ua);
    // entropy wave u
    alpha = (r[0] - l[0]) - dp / aa; s = 0.5 * alpha * std::abs(u);
    flux[0] -= s; flux[1] -= s * u; flux[2] -= s * v; flux[3] -= s * w; flux[4] -= s * ke;
    // shear waves u
    alpha = rho * (r[2] - l[2]); s = 0.5 * alpha * std::abs(u);
    flux[2] -= s; flux[4] -= s * v;
    alpha = rho * (r[3] - l[3]); s = 0.5 * alpha * std::abs(u);
    flux[3] -= s; flux[4] -= s * w;
    // acoustic wave u + a
    alpha = (dp + rho * a * du) / (2 * aa); s = 0.5 * alpha * std::abs(u + a);
    flux[0] -= s; flux[1] -= s * (u + a); flux[2] -= s * v; flux[3] -= s * w; flux[4] -= s * (h + ua);
    return flux;
  }
};

enum RiemannKind { kExact = 0, kHllc = 1, kRoe = 2 };

// --------------------------------------------------------- rotated wrapper
// riemann/rotated/euler.hpp:24-356: frame = {normal a, tangent b, bitangent c}.
struct Frame { Vec3 a, b, c; };

inline void GlobalToNormal(Frame const &f, Val5 *v) {  // euler.hpp:117-125
  Vec3 p{ (*v)[1], (*v)[2], (*v)[3] };
  (*v)[1] = dot(p, f.a); (*v)[2] = dot(p, f.b); (*v)[3] = dot(p, f.c);
}
inline void NormalToGlobal(Frame const &f, Val5 *v) {  // euler.hpp:126-134
  double pa = (*v)[1], pb = (*v)[2], pc = (*v)[3];
  (*v)[1] = pa * f.a[0] + pb * f.b[0] + pc * f.c[0];
  (*v)[2] = pa * f.a[1] + pb * f.b[1] + pc * f.c[1];
  (*v)[3] = pa * f.a[2] + pb * f.b[2] + pc * f.c[2];
}

struct RotatedEuler {
  RiemannKind kind;
  ExactSolver exact;  // one per face point, as riemann_[face][q] in fem.hpp:76-77
  Val5 NormalPairToGlobalFlux(Frame const &f, Val5 const &l, Val5 const &r) {
    Val5 flux;
    switch (kind) {
    case kExact: flux = exact.FluxUpwind(l, r); break;
    case kHllc: flux = HllcSolver().FluxUpwind(l, r); break;
    default: flux = RoeSolver().FluxUpwind(l, r); break;
    }
    NormalToGlobal(f, &flux);
    return flux;
  }
  // euler.hpp:153-159
  Val5 GetFluxUpwind(Frame const &f, Val5 const &cl, Val5 const &cr) {
    Val5 l = ConservativeToPrimitive(cl), r = ConservativeToPrimitive(cr);
    GlobalToNormal(f, &l); GlobalToNormal(f, &r);
    return NormalPairToGlobalFlux(f, l, r);
  }
  // euler.hpp:160-174, the default branch (no reflected Riemann problem).
  Val5 GetFluxOnInviscidWall(Frame const &f, Val5 const &c) {
    Val5 pv = ConservativeToPrimitive(c);
    Val5 flux{ 0, pv[4], 0, 0, 0 };
    NormalToGlobal(f, &flux);
    return flux;
  }
  // euler.hpp:144-151: supersonic in/outlet = physical normal flux.
  Val5 GetFluxOfState(Frame const &f, Val5 const &c) {
    Val5 pv = ConservativeToPrimitive(c);
    GlobalToNormal(f, &pv);
    Val5 flux = PrimitiveToFlux(pv);
    NormalToGlobal(f, &flux);
    return flux;
  }
  // euler.hpp:296-299
  Val5 GetFluxOnSmartBoundary(Frame const &f, Val5 const &ci, Val5 const &ce) {
    return GetFluxUpwind(f, ci, ce);
  }
  // euler.hpp:181-246, the default branch (no boundary Riemann problem):
  // given = (p_total, u_cos, v_cos, w_cos, T_total).
  Val5 GetFluxOnSubsonicInlet(Frame const &f, Val5 const &ci, Val5 const &given) {
    constexpr double kR = 287.05, kGm1 = 0.4, kGamma = 1.4;  // types.hpp:206-208
    Val5 pi = ConservativeToPrimitive(ci);
    double c_interior = (pi[0] > 0 && pi[4] > 0) ? std::sqrt(kGamma * pi[4] / pi[0]) : 0;
    GlobalToNormal(f, &pi);
    double riemann_interior = -pi[1] - c_interior * (2 / kGm1);
    double u_cos = given[1], v_cos = given[2], w_cos = given[3], t_total = given[4];
    double a = f.a[0] * u_cos + f.a[1] * v_cos + f.a[2] * w_cos;
    double a2 = a * a, gamma_times_r = kGamma * kR;
    auto larger = [](double qa, double b_half, double qc) {
      double sqrt_delta = std::sqrt(b_half * b_half - qa * qc);
      double x1 = (-b_half + sqrt_delta) / qa, x2 = (-b_half - sqrt_delta) / qa;
      return std::max(x1, x2);
    };
    double c_boundary = larger(a2 + 2 / kGm1, riemann_interior,
        kGm1 / 2 * (riemann_interior * riemann_interior) - a2 * gamma_times_r * t_total);
    double t_boundary = c_boundary * c_boundary / gamma_times_r;
    double mach = std::sqrt((t_total / t_boundary - 1) / (kGm1 / 2));  // types.hpp:273-280
    double p_boundary = given[0] / std::pow(1 + kGm1 / 2 * mach * mach, kGamma / kGm1);  // :268-272
    double speed = mach * c_boundary;
    Val5 pb{ p_boundary / kR / t_boundary, u_cos * speed, v_cos * speed, w_cos * speed, p_boundary };
    GlobalToNormal(f, &pb);
    Val5 flux = PrimitiveToFlux(pb);
    NormalToGlobal(f, &flux);
    return flux;
  }
  // euler.hpp:263-280, the default branch: given[4] = back pressure.
  Val5 GetFluxOnSubsonicOutlet(Frame const &f, Val5 const &ci, Val5 const &given) {
    Val5 pb = ConservativeToPrimitive(ci);
    double rt_interior = pb[4] / pb[0];
    pb[4] = given[4];
    pb[0] = pb[4] / rt_interior;
    GlobalToNormal(f, &pb);
    Val5 flux = PrimitiveToFlux(pb);
    NormalToGlobal(f, &flux);
    return flux;
  }
};

}  // namespace oracle

#endif  // ORACLE_CORE_HPP_
