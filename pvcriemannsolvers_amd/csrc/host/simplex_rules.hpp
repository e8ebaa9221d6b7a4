// Symmetric simplex quadrature rules of the tetrahedral DG path.
//
// The constants solve the moment equations of each rule (derivation:
// derive_simplex_rules.py next to this file); the orbit structure and point
// order are those of the reference's integrator/tetrahedron.hpp:150-187
// (14 points, exact to degree 5) and integrator/triangle.hpp:163-188
// (6 points, exact to degree 4).  Local weights sum to the measure of the
// reference simplex (1/6 and 1/2).
#ifndef MINICFD_HOST_SIMPLEX_RULES_HPP_
#define MINICFD_HOST_SIMPLEX_RULES_HPP_

namespace mcfd {

// points (xi, eta, zeta) and weights of the 14-point tetrahedron rule
inline void TetRule14(double x[14][3], double w[14]) {
  const double a31[2] = { 0.3108859192633006098, 0.092735250310891226402 };
  const double w31[2] = { 0.0187813209530026418, 0.012248840519393658257 };
  int q = 0;
  for (int o = 0; o < 2; ++o) {
    const double a = a31[o], c = 1 - 3 * a;
    const double pts[4][3] = { { a, a, a }, { a, a, c }, { a, c, a }, { c, a, a } };
    for (int i = 0; i < 4; ++i, ++q) {
      for (int d = 0; d < 3; ++d) x[q][d] = pts[i][d];
      w[q] = w31[o];
    }
  }
  const double b = 0.045503704125649649492, c = (1 - 2 * b) * 0.5;
  const double pts[6][3] = { { b, b, c }, { b, c, b }, { b, c, c }, { c, b, b }, { c, b, c }, { c, c, b } };
  for (int i = 0; i < 6; ++i, ++q) {
    for (int d = 0; d < 3; ++d) x[q][d] = pts[i][d];
    w[q] = 0.007091003462846911073;
  }
}

// points (a, b) and weights of the 6-point triangle rule
inline void TriRule6(double x[6][2], double w[6]) {
  const double a21[2] = { 0.44594849091596488632, 0.09157621350977074346 };
  const double w21[2] = { 0.11169079483900573285, 0.054975871827660933819 };
  int q = 0;
  for (int o = 0; o < 2; ++o) {
    const double a = a21[o], b = 1 - a - a;
    const double pts[3][2] = { { a, a }, { a, b }, { b, a } };
    for (int i = 0; i < 3; ++i, ++q) {
      x[q][0] = pts[i][0]; x[q][1] = pts[i][1];
      w[q] = w21[o];
    }
  }
}

}  // namespace mcfd

#endif  // MINICFD_HOST_SIMPLEX_RULES_HPP_
