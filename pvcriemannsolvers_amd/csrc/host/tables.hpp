// 1-D quadrature rules, Lagrange derivative matrices and ESFR correction
// derivatives used to fill the kernels' LineTables.
//
//   Gauss-Legendre points/weights  integrator/legendre.hpp:36-108 (closed forms)
//   Gauss-Lobatto points/weights   integrator/lobatto.hpp:37-80
//   Lagrange basis on the points   basis/lagrange.hpp:34-145
//   Vincent g_L', g_R'             basis/vincent.hpp:19-102
#ifndef MINICFD_HOST_TABLES_HPP_
#define MINICFD_HOST_TABLES_HPP_

#include <cmath>
#include <stdexcept>

namespace mcfd {

inline void GaussRule(bool lobatto, int q, double *x, double *w) {
  if (!lobatto) {
    switch (q) {
    case 1: x[0] = 0; w[0] = 2; return;
    case 2: x[1] = std::sqrt(1.0 / 3.0); x[0] = -x[1]; w[0] = w[1] = 1; return;
    case 3: x[2] = std::sqrt(0.6); x[1] = 0; x[0] = -x[2];
            w[0] = w[2] = 5.0 / 9.0; w[1] = 8.0 / 9.0; return;
    case 4: {
      double o = std::sqrt((3 + 2 * std::sqrt(1.2)) / 7), i = std::sqrt((3 - 2 * std::sqrt(1.2)) / 7);
      x[0] = -o; x[1] = -i; x[2] = i; x[3] = o;
      w[0] = w[3] = (18 - std::sqrt(30.0)) / 36;
      w[1] = w[2] = (18 + std::sqrt(30.0)) / 36;
      return;
    }
    case 5: {
      double o = std::sqrt((5 + std::sqrt(40 / 7.0)) / 9), i = std::sqrt((5 - std::sqrt(40 / 7.0)) / 9);
      x[0] = -o; x[1] = -i; x[2] = 0; x[3] = i; x[4] = o;
      w[0] = w[4] = (322 - 13 * std::sqrt(70.0)) / 900;
      w[1] = w[3] = (322 + 13 * std::sqrt(70.0)) / 900;
      w[2] = 128.0 / 225.0;
      return;
    }
    case 6: {
      const double p[3] = { 0.932469514203152, 0.6612093864662645, 0.23861918608319693 };
      const double v[3] = { 0.17132449237917016, 0.36076157304813855, 0.4679139345726912 };
      for (int k = 0; k < 3; ++k) { x[k] = -p[k]; x[5 - k] = p[k]; w[k] = w[5 - k] = v[k]; }
      return;
    }
    }
  } else {
    switch (q) {
    case 2: x[0] = -1; x[1] = 1; w[0] = w[1] = 1; return;
    case 3: x[0] = -1; x[1] = 0; x[2] = 1; w[0] = w[2] = 1.0 / 3; w[1] = 4.0 / 3; return;
    case 4: x[0] = -1; x[2] = std::sqrt(1.0 / 5); x[1] = -x[2]; x[3] = 1;
            w[0] = w[3] = 1.0 / 6; w[1] = w[2] = 5.0 / 6; return;
    case 5: x[0] = -1; x[3] = std::sqrt(21.0) / 7; x[1] = -x[3]; x[2] = 0; x[4] = 1;
            w[0] = w[4] = 0.1; w[1] = w[3] = 49.0 / 90; w[2] = 32.0 / 45; return;
    case 6: {
      double o = std::sqrt((7 + 2 * std::sqrt(7.0)) / 21), i = std::sqrt((7 - 2 * std::sqrt(7.0)) / 21);
      x[0] = -1; x[1] = -o; x[2] = -i; x[3] = i; x[4] = o; x[5] = 1;
      w[0] = w[5] = 1.0 / 15;
      w[1] = w[4] = (14.0 - std::sqrt(7.0)) / 30;
      w[2] = w[3] = (14.0 + std::sqrt(7.0)) / 30;
      return;
    }
    }
  }
  throw std::invalid_argument("unsupported quadrature size");
}

// Lagrange basis function m on points x[0..q) evaluated at s, and its derivative.
inline double LagrangeValue(int q, const double *x, int m, double s) {
  double v = 1;
  for (int l = 0; l < q; ++l) if (l != m) v *= (s - x[l]) / (x[m] - x[l]);
  return v;
}
inline double LagrangeDerivative(int q, const double *x, int m, double s) {
  double sum = 0;
  for (int l = 0; l < q; ++l) {
    if (l == m) continue;
    double t = 1 / (x[m] - x[l]);
    for (int o = 0; o < q; ++o) if (o != m && o != l) t *= (s - x[o]) / (x[m] - x[o]);
    sum += t;
  }
  return sum;
}

// Legendre polynomial P_k(s) and derivative by the three-term recurrence.
inline void Legendre(int k, double s, double *p, double *dp) {
  double p0 = 1, p1 = s, d0 = 0, d1 = 1;
  if (k == 0) { *p = 1; *dp = 0; return; }
  for (int n = 1; n < k; ++n) {
    double p2 = ((2 * n + 1) * s * p1 - n * p0) / (n + 1);
    double d2 = d0 + (2 * n + 1) * p1;  // P'_{n+1} = P'_{n-1} + (2n+1) P_n
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  *p = p1; *dp = d1;
}

// d/ds of g_R(s) = (P_k + c_prev P_{k-1} + c_next P_{k+1}) / 2, c_prev = 1 - c_next.
inline double CorrectionRightDerivative(int k, double c_next, double s) {
  double p, dk, dkm, dkp;
  Legendre(k, s, &p, &dk);
  Legendre(k - 1, s, &p, &dkm);
  Legendre(k + 1, s, &p, &dkp);
  return 0.5 * (dk + (1 - c_next) * dkm + c_next * dkp);
}
inline double CorrectionLeftDerivative(int k, double c_next, double s) {
  return -CorrectionRightDerivative(k, c_next, -s);
}

}  // namespace mcfd

#endif  // MINICFD_HOST_TABLES_HPP_
