// Oracle (test infrastructure): symmetric simplex quadrature rules.
//
// Orbit parameters and weights solve the moment equations of their degree
// (derivation: pvcriemannsolvers_amd/csrc/host/derive_simplex_rules.py); the
// orbit structure and point order are those of the reference's
// integrator/tetrahedron.hpp:150-187 (14 points, degree 5) and
// integrator/triangle.hpp:163-188 (6 points, degree 4).  Local weights sum to
// the reference simplex measure (1/6 and 1/2).  Pinned by monomial exactness
// and by the reference's digits (tests/test_oracle_dg.py).
#ifndef ORACLE_SIMPLEX_HPP_
#define ORACLE_SIMPLEX_HPP_

#include <array>

namespace oracle {

struct TetRule14 {
  std::array<std::array<double, 3>, 14> x;
  std::array<double, 14> w;
  TetRule14() {
    const double a31[2] = { 0.3108859192633006098, 0.092735250310891226402 };
    const double w31[2] = { 0.0187813209530026418, 0.012248840519393658257 };
    int q = 0;
    for (int o = 0; o < 2; ++o) {
      double a = a31[o], c = 1 - 3 * a;
      x[q++] = { a, a, a }; x[q++] = { a, a, c }; x[q++] = { a, c, a }; x[q++] = { c, a, a };
      for (int i = 0; i < 4; ++i) w[4 * o + i] = w31[o];
    }
    double b = 0.045503704125649649492, c = (1 - 2 * b) * 0.5;
    x[q++] = { b, b, c }; x[q++] = { b, c, b }; x[q++] = { b, c, c };
    x[q++] = { c, b, b }; x[q++] = { c, b, c }; x[q++] = { c, c, b };
    for (int i = 8; i < 14; ++i) w[i] = 0.007091003462846911073;
  }
};

struct TriRule6 {
  std::array<std::array<double, 2>, 6> x;
  std::array<double, 6> w;
  TriRule6() {
    const double a21[2] = { 0.44594849091596488632, 0.09157621350977074346 };
    const double w21[2] = { 0.11169079483900573285, 0.054975871827660933819 };
    int q = 0;
    for (int o = 0; o < 2; ++o) {
      double a = a21[o], b = 1 - a - a;
      x[q++] = { a, a }; x[q++] = { a, b }; x[q++] = { b, a };
      for (int i = 0; i < 3; ++i) w[3 * o + i] = w21[o];
    }
  }
};

}  // namespace oracle

#endif  // ORACLE_SIMPLEX_HPP_
