// Oracle (test infrastructure): modal DG on hexahedra and tetrahedra,
// restated from spatial/dg/general.hpp:51-202 over spatial/fem.hpp:224-435,
// polynomial/projection.hpp:28-164, basis/linear.hpp:87-141,
// basis/taylor.hpp:198-208,394-429 and integrator/function.hpp:139-168.
//
// Each cell carries an orthonormal basis phi = S T(x - c): T the Taylor
// monomials of degree <= p in global coordinates, c the cell centre and S
// the lower-triangular Gram-Schmidt factor built from the cell's own
// quadrature.  The state is the K x N coefficient block of every cell,
// K-contiguous per basis function (expansion.hpp:83-104); no mass matrix is
// needed because the basis is orthonormal.
//
// Pinned by (tests/test_oracle_dg.py): the simplex rules' monomial exactness
// and the reference's digits; orthonormality of each cell's basis under its
// quadrature and exact projection of polynomials of degree <= p; the
// closed-form residual R_i = -(div F) sqrt(V) delta_i0 of a linear pressure
// field with constant density and velocity (flux-of-own-state boundaries);
// the first-order finite-volume update that piecewise-constant data must
// give, recomputed face by face; free-stream preservation, conservation
// with walls, and the sampled time step.
#ifndef ORACLE_DG_HPP_
#define ORACLE_DG_HPP_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core.hpp"
#include "fr.hpp"
#include "mesh.hpp"

namespace oracle {

// basis/taylor.hpp:198-208 (degree 1 and 2 in three dimensions).
inline int TaylorSize(int p) { return p == 1 ? 4 : 10; }
inline void TaylorValues(int p, Vec3 const &x, double *t) {
  t[0] = 1; t[1] = x[0]; t[2] = x[1]; t[3] = x[2];
  if (p >= 2) {
    t[4] = x[0] * x[0]; t[5] = x[0] * x[1]; t[6] = x[0] * x[2];
    t[7] = x[1] * x[1]; t[8] = x[1] * x[2]; t[9] = x[2] * x[2];
  }
}

struct ModalBasis {  // basis::OrthoNormal
  int p = 2, n = 10;
  Vec3 center{};
  double s[10][10] = {};  // lower triangular

  // integrator/function.hpp:139-168 (OrthoNormalize) on the cell quadrature.
  void Build(int degree, Vec3 const &c, std::vector<Vec3> const &qp, std::vector<double> const &qw) {
    p = degree; n = TaylorSize(p); center = c;
    double a[10][10] = {};
    for (size_t q = 0; q < qp.size(); ++q) {
      double t[10];
      TaylorValues(p, sub(qp[q], center), t);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) a[i][j] += t[i] * t[j] * qw[q];
    }
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) s[i][j] = (i == j);
    s[0][0] = 1 / std::sqrt(a[0][0]);
    for (int i = 1; i < n; ++i) {
      for (int j = 0; j < i; ++j) {
        double temp = 0;
        for (int k = 0; k <= j; ++k) temp += s[j][k] * a[k][i];
        for (int l = 0; l <= j; ++l) s[i][l] -= temp * s[j][l];
      }
      double norm_sq = 0;
      for (int j = 0; j <= i; ++j) {
        double sum = 0, sij = s[i][j];
        for (int k = 0; k < j; ++k) sum += 2 * s[i][k] * a[k][j];
        norm_sq += sij * (sij * a[j][j] + sum);
      }
      double inv = std::sqrt(norm_sq);
      for (int j = 0; j <= i; ++j) s[i][j] /= inv;
    }
  }
  // OrthoNormal::operator(): S T(x - c).
  void Values(Vec3 const &x, double *phi) const {
    double t[10];
    TaylorValues(p, sub(x, center), t);
    for (int i = 0; i < n; ++i) {
      double v = 0;
      for (int j = 0; j <= i; ++j) v += s[i][j] * t[j];
      phi[i] = v;
    }
  }
  // Taylor::GetGradValue (taylor.hpp:394-429): grad[i][d] = d phi_i / d x_d.
  void Grads(Vec3 const &xg, double (*g)[3]) const {
    Vec3 x = sub(xg, center);
    for (int i = 0; i < n; ++i) {
      g[i][0] = s[i][1]; g[i][1] = s[i][2]; g[i][2] = s[i][3];
      if (p >= 2) {
        g[i][0] += s[i][4] * (2 * x[0]); g[i][0] += s[i][5] * x[1]; g[i][0] += s[i][6] * x[2];
        g[i][1] += s[i][5] * x[0]; g[i][1] += s[i][7] * (2 * x[1]); g[i][1] += s[i][8] * x[2];
        g[i][2] += s[i][6] * x[0]; g[i][2] += s[i][8] * x[1]; g[i][2] += s[i][9] * (2 * x[2]);
      }
    }
  }
};

struct DgScheme {
  Mesh const *mesh;
  int p, n;
  RotatedEuler riemann;                 // the solvers that keep no state
  std::vector<RotatedEuler> riemanns;   // [face][point], fem.hpp:95-115 (the exact solver's star_u)
  std::vector<BcEntry> bcs;
  // Given states per boundary face point, [boundary face][point] (the BC
  // functions f(x, t) at the residual's time); empty: each tag's constant.
  std::vector<Val5> point_state;
  std::vector<ModalBasis> basis;
  // Cell::adj_cells_ (mesh/part.hpp:962-963): neighbours in interior-face order.
  std::vector<std::vector<int64_t>> adj;
  std::vector<std::vector<int64_t>> adj_faces;  // Cell::adj_faces_ (part.hpp:971-972)
  // Limiter applied at every SetSolutionColumn (spatial/with_limiter.hpp:56-59):
  // 0 none; 1 weno::Lazy reconstruction of the cells weno::Eigen::IsNotSmooth
  // flags; 2 weno::Lazy on every cell (Lazy::IsNotSmooth is always true);
  // 3 weno::Eigen (characteristic-wise) reconstruction of the flagged cells;
  // 4 weno::Eigen reconstruction of every cell.
  int limiter = 0;
  double weno_w0 = 0.001, weno_eps = 1e-6;
  std::vector<uint8_t> flags;  // of the last limiter application

  // index of point q of face fi in [face][point] order
  size_t FacePoint(size_t fi, size_t q) const { return fi * mesh->faces[0].qp.size() + q; }

  DgScheme(Mesh const *m, int degree, int riemann_kind, std::vector<BcEntry> bc)
      : mesh(m), p(degree), n(TaylorSize(degree)), bcs(std::move(bc)) {
    if (degree < 1 || degree > 2) throw std::runtime_error("modal DG: degree must be 1 or 2");
    riemann.kind = RiemannKind(riemann_kind);
    for (auto const &face : mesh->faces)
      for (size_t q = 0; q < face.qp.size(); ++q) riemanns.push_back(riemann);
    basis.resize(mesh->cells.size());
    for (size_t c = 0; c < mesh->cells.size(); ++c) {
      auto const &cell = mesh->cells[c];
      basis[c].Build(p, cell.center, cell.qp, cell.qw);
    }
    adj.resize(mesh->cells.size());
    adj_faces.resize(mesh->cells.size());
    for (int64_t f = 0; f < mesh->n_interior; ++f) {
      adj[mesh->faces[f].holder].push_back(mesh->faces[f].sharer);
      adj[mesh->faces[f].sharer].push_back(mesh->faces[f].holder);
      adj_faces[mesh->faces[f].holder].push_back(f);
      adj_faces[mesh->faces[f].sharer].push_back(f);
    }
  }
  size_t Stride() const { return size_t(5) * n; }

  // Projection::GlobalToValue: u = C phi(x).
  Val5 Value(int64_t c, double const *u, Vec3 const &x) const {
    double phi[10];
    basis[c].Values(x, phi);
    Val5 v{ 0, 0, 0, 0, 0 };
    double const *cf = u + c * Stride();
    for (int k = 0; k < 5; ++k) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += cf[5 * i + k] * phi[i];
      v[k] = s;
    }
    return v;
  }
  // Polynomial::AddToResidual of flux * phi(x)^T.
  void AddRankOne(int64_t c, Val5 const &flux, Vec3 const &x, double *r) const {
    double phi[10];
    basis[c].Values(x, phi);
    double *rc = r + c * Stride();
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 5; ++k) rc[5 * i + k] += flux[k] * phi[i];
  }

  // fem.hpp:237-247 (GetResidualColumn) with dg/general.hpp:58-89, 92-202.
  void Residual(double const *u, double *r) {
    std::fill(r, r + mesh->cells.size() * Stride(), 0.0);
    for (size_t c = 0; c < mesh->cells.size(); ++c) {  // AddFluxDivergence
      auto const &cell = mesh->cells[c];
      double *rc = r + c * Stride();
      for (size_t q = 0; q < cell.qp.size(); ++q) {
        Val5 v = Value(c, u, cell.qp[q]);
        Flux53 f = GetFluxMatrix(v);
        for (auto &row : f) for (auto &x : row) x *= cell.qw[q];
        double g[10][3];
        basis[c].Grads(cell.qp[q], g);
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < 5; ++k) rc[5 * i + k] += f[k][0] * g[i][0] + f[k][1] * g[i][1] + f[k][2] * g[i][2];
      }
    }
    for (int64_t fi = 0; fi < mesh->n_interior; ++fi) {  // AddFluxToHolderAndSharer
      auto const &face = mesh->faces[fi];
      for (size_t q = 0; q < face.qp.size(); ++q) {
        Vec3 xs = add(face.qp[q], face.shift);
        Val5 uh = Value(face.holder, u, face.qp[q]);
        Val5 us = Value(face.sharer, u, xs);
        Val5 flux = riemanns[FacePoint(fi, q)].GetFluxUpwind(face.frames[q], uh, us);
        for (auto &x : flux) x *= -face.qw[q];
        AddRankOne(face.holder, flux, face.qp[q], r);
        for (auto &x : flux) x = -x;
        AddRankOne(face.sharer, flux, xs, r);
      }
    }
    // fem.hpp:261-323: supersonic inlets and outlets, subsonic inlets and outlets, walls, smart.
    for (int family : kBcFamilyOrder) {
      for (size_t fi = mesh->n_interior; fi < mesh->faces.size(); ++fi) {
        auto const &face = mesh->faces[fi];
        if (face.tag < 0 || face.tag >= (int)bcs.size()) throw std::runtime_error("boundary tag without BC");
        if (bcs[face.tag].kind != family) continue;
        for (size_t q = 0; q < face.qp.size(); ++q) {
          Val5 flux;
          auto const &frame = face.frames[q];
          auto &rs = riemanns[FacePoint(fi, q)];
          Val5 const &given = point_state.empty()
              ? bcs[face.tag].state : point_state[FacePoint(fi, q) - FacePoint(mesh->n_interior, 0)];
          if (family == kBcSupersonicInlet) {
            flux = rs.GetFluxOfState(frame, given);
          } else {
            Val5 uh = Value(face.holder, u, face.qp[q]);
            if (family == kBcSupersonicOutlet) flux = rs.GetFluxOfState(frame, uh);
            else if (family == kBcInviscidWall) flux = rs.GetFluxOnInviscidWall(frame, uh);
            else if (family == kBcSubsonicInlet) flux = rs.GetFluxOnSubsonicInlet(frame, uh, given);
            else if (family == kBcSubsonicOutlet) flux = rs.GetFluxOnSubsonicOutlet(frame, uh, given);
            else flux = rs.GetFluxOnSmartBoundary(frame, uh, given);
          }
          for (auto &x : flux) x *= -face.qw[q];
          AddRankOne(face.holder, flux, face.qp[q], r);
        }
      }
    }
  }

  // projection.hpp:28-43 (Project): coefficients of f given at the cell
  // quadrature points, values[cell][q][5].
  void Approximate(double const *values, double *u) const {
    size_t off = 0;
    for (size_t c = 0; c < mesh->cells.size(); ++c) {
      auto const &cell = mesh->cells[c];
      double *cf = u + c * Stride();
      std::fill(cf, cf + Stride(), 0.0);
      for (size_t q = 0; q < cell.qp.size(); ++q, off += 5) {
        double phi[10];
        basis[c].Values(cell.qp[q], phi);
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < 5; ++k) {
            double prod = values[off + k] * phi[i];
            prod *= cell.qw[q];
            cf[5 * i + k] += prod;
          }
      }
    }
  }

  // ---------------------------------------------------------------- WENO limiter
  // Projection::average (projection.hpp:45-50): C[:, 0] * S(0, 0).
  Val5 Average(int64_t c, double const *coeff) const {
    Val5 v;
    for (int k = 0; k < 5; ++k) v[k] = coeff[k] * basis[c].s[0][0];
    return v;
  }
  // limiter/weno.hpp:230-258 (weno::Eigen::IsNotSmooth) on rho and E.
  bool IsNotSmooth(int64_t c, double const *u) const {
    const int comps[2] = { 0, 4 };
    Val5 max_abs = Average(c, u + c * Stride());
    for (int i : comps) max_abs[i] = std::max(1e-9, std::abs(max_abs[i]));
    double sum_abs[5] = { 0, 0, 0, 0, 0 };
    const Vec3 &x = mesh->cells[c].center;
    Val5 mine = Value(c, u, x);
    for (int64_t a : adj[c]) {
      Val5 theirs = Value(a, u, x);
      Val5 avg = Average(a, u + a * Stride());
      for (int i : comps) {
        sum_abs[i] += std::abs(mine[i] - theirs[i]);
        max_abs[i] = std::max(max_abs[i], std::abs(avg[i]));
      }
    }
    const double volume_power = (p + 1.0) / 2.0 / 3;
    double divisor = std::pow(mesh->cells[c].volume, volume_power);
    divisor *= adj[c].size();
    const double reference = p < 3 ? 1.0 : 3.0;
    for (int i : comps)
      if (sum_abs[i] / max_abs[i] / divisor > reference) return true;
    return false;
  }
  // weno.hpp:74-95 (GetSmoothness) of coefficients on cell c's basis.
  Val5 Smoothness(int64_t c, double const *coeff) const {
    auto const &b = basis[c];
    double tc[5][10] = {};  // Taylor coefficients C S
    for (int k = 0; k < 5; ++k)
      for (int j = 0; j < n; ++j) {
        double v = 0;
        for (int i = 0; i < n; ++i) v += coeff[5 * i + k] * b.s[i][j];
        tc[k][j] = v;
      }
    auto const &cell = mesh->cells[c];
    double integral[5][10] = {};
    for (size_t q = 0; q < cell.qp.size(); ++q) {
      Vec3 y = sub(cell.qp[q], b.center);
      for (int k = 0; k < 5; ++k) {
        double pdv[10];  // taylor.hpp:241-281 (GetPartialDerivatives)
        pdv[0] = 0;
        for (int j = 1; j < n; ++j) pdv[j] = tc[k][j];
        if (p >= 2) {
          pdv[1] += tc[k][4] * (2 * y[0]); pdv[1] += tc[k][5] * y[1]; pdv[1] += tc[k][6] * y[2];
          pdv[2] += tc[k][5] * y[0]; pdv[2] += tc[k][7] * (2 * y[1]); pdv[2] += tc[k][8] * y[2];
          pdv[3] += tc[k][6] * y[0]; pdv[3] += tc[k][8] * y[1]; pdv[3] += tc[k][9] * (2 * y[2]);
          pdv[4] += tc[k][4]; pdv[7] += tc[k][7]; pdv[9] += tc[k][9];
        }
        for (int j = 0; j < n; ++j) integral[k][j] += pdv[j] * pdv[j] * cell.qw[q];
      }
    }
    const double volume = 1 / (b.s[0][0] * b.s[0][0]);  // OrthoNormal::Measure
    auto weight = [volume](int k) {
      double f = std::tgamma(k + 1);
      return std::pow(volume, k * 2. / 3 - 1) / (f * f);
    };
    Val5 beta;
    for (int k = 0; k < 5; ++k) {
      double s1 = integral[k][1];
      s1 += integral[k][2];
      s1 += integral[k][3];
      s1 *= weight(1);
      if (p >= 2) {
        double s2 = integral[k][4];
        for (int j = 5; j < 10; ++j) s2 += integral[k][j];
        s2 *= weight(2);
        s1 += s2;
      }
      beta[k] = s1;
    }
    return beta;
  }
  // weno.hpp:121-155 (Borrow): neighbour expansions projected onto cell c's
  // basis with their averages moved onto c's, then c's own expansion last.
  std::vector<std::vector<double>> Borrow(int64_t c, double const *u) const {
    auto const &cell = mesh->cells[c];
    auto const &b = basis[c];
    const int n_adj = (int)adj[c].size();
    std::vector<std::vector<double>> proj(n_adj + 1, std::vector<double>(Stride(), 0.0));
    Val5 my_avg = Average(c, u + c * Stride());
    for (int ia = 0; ia < n_adj; ++ia) {
      auto &pr = proj[ia];
      for (size_t q = 0; q < cell.qp.size(); ++q) {  // Projection::Approximate of the neighbour
        Val5 f = Value(adj[c][ia], u, cell.qp[q]);
        double phi[10];
        b.Values(cell.qp[q], phi);
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < 5; ++k) {
            double prod = f[k] * phi[i];
            prod *= cell.qw[q];
            pr[5 * i + k] += prod;
          }
      }
      Val5 its_avg = Average(c, pr.data());
      for (int k = 0; k < 5; ++k) {  // ProjectionWrapper::operator+=(Value)
        double offset = my_avg[k] - its_avg[k];
        offset /= b.s[0][0];
        pr[k] += offset;
      }
    }
    std::copy(u + c * Stride(), u + (c + 1) * Stride(), proj[n_adj].begin());
    return proj;
  }
  // Nonlinear weights (weno.hpp:160-177): w0 for neighbours, 1 - n w0 for the
  // cell itself, divided by (beta + eps)^2 and normalised.
  std::vector<Val5> Weights(int64_t c, std::vector<std::vector<double>> const &proj) const {
    const int n_adj = (int)proj.size() - 1;
    std::vector<Val5> weights(n_adj + 1);
    for (int ia = 0; ia <= n_adj; ++ia)
      for (int k = 0; k < 5; ++k) weights[ia][k] = weno_w0;
    for (int k = 0; k < 5; ++k) { weights[n_adj][k] *= -n_adj; weights[n_adj][k] += 1.0; }
    for (int ia = 0; ia <= n_adj; ++ia) {
      Val5 beta = Smoothness(c, proj[ia].data());
      for (int k = 0; k < 5; ++k) {
        beta[k] += weno_eps;
        beta[k] *= beta[k];
        weights[ia][k] /= beta[k];
      }
    }
    Val5 sum{ 0, 0, 0, 0, 0 };
    for (auto const &w : weights) for (int k = 0; k < 5; ++k) sum[k] += w[k];
    for (auto &w : weights) for (int k = 0; k < 5; ++k) w[k] /= sum[k];
    return weights;
  }
  // new = w_self proj_self + sum_i w_i proj_i (weno.hpp:178-194).
  void Blend(std::vector<std::vector<double>> const &proj, std::vector<Val5> const &w, double *out) const {
    const int n_adj = (int)proj.size() - 1;
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 5; ++k) {
        double v = proj[n_adj][5 * i + k] * w[n_adj][k];
        for (int ia = 0; ia < n_adj; ++ia) v += proj[ia][5 * i + k] * w[ia][k];
        out[5 * i + k] = v;
      }
  }
  // weno.hpp:121-195 (weno::Lazy::Borrow + Reconstruct) of cell c from u.
  void LazyReconstruct(int64_t c, double const *u, double *out) const {
    auto proj = Borrow(c, u);
    Blend(proj, Weights(c, proj), out);
  }
  // riemann/euler/eigen.hpp:26-56 (EigenMatrices) of the state `cv` in the
  // frame (nu, mu, pi); L and R row-major 5 x 5.
  static void EigenMatrices(Val5 const &cv, Frame const &f, double l[5][5], double r[5][5]) {
    Val5 pv = ConservativeToPrimitive(cv);
    const double ux = pv[1], uy = pv[2], uz = pv[3];
    const Vec3 &nu = f.a, &mu = f.b, &pi = f.c;
    const double u_nu = ux * nu[0] + uy * nu[1] + uz * nu[2];
    const double u_mu = ux * mu[0] + uy * mu[1] + uz * mu[2];
    const double u_pi = ux * pi[0] + uy * pi[1] + uz * pi[2];
    const double a = std::sqrt(1.4 * pv[4] / pv[0]);
    const double ek = (ux * ux + uy * uy + uz * uz) / 2;
    const double e0 = a * a * (1 / (1.4 - 1));
    const double h0 = e0 + ek;
    const double rc[5][5] = {  // columns of R
      { 1, ux - a * nu[0], uy - a * nu[1], uz - a * nu[2], h0 - a * u_nu },
      { 1, ux, uy, uz, ek },
      { 0, mu[0], mu[1], mu[2], u_mu },
      { 0, pi[0], pi[1], pi[2], u_pi },
      { 1, ux + a * nu[0], uy + a * nu[1], uz + a * nu[2], h0 + a * u_nu },
    };
    for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) r[i][j] = rc[j][i];
    const double b1 = 1 / e0, b2 = b1 * ek;
    const double lc[5][5] = {  // columns of L
      { (b2 + u_nu / a) / 2, 1 - b2, -u_mu, -u_pi, (b2 - u_nu / a) / 2 },
      { -(b1 * ux + nu[0] / a) / 2, b1 * ux, mu[0], pi[0], -(b1 * ux - nu[0] / a) / 2 },
      { -(b1 * uy + nu[1] / a) / 2, b1 * uy, mu[1], pi[1], -(b1 * uy - nu[1] / a) / 2 },
      { -(b1 * uz + nu[2] / a) / 2, b1 * uz, mu[2], pi[2], -(b1 * uz - nu[2] / a) / 2 },
      { b1 / 2, -b1, 0, 0, b1 / 2 },
    };
    for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) l[i][j] = lc[j][i];
  }
  // weno.hpp:292-348 (weno::Eigen::ReconstructOnFace + Reconstruct): per
  // adjacent face, WENO in the characteristic variables of the cell average
  // in that face's frame, mapped back and averaged with the neighbours'
  // volumes as weights.
  void EigenReconstruct(int64_t c, double const *u, double *out) const {
    auto proj = Borrow(c, u);
    const Val5 big_u = Average(c, u + c * Stride());
    std::vector<double> acc(Stride(), 0.0), tmp(Stride());
    double total_volume = 0;
    for (size_t i_face = 0; i_face < adj_faces[c].size(); ++i_face) {
      auto const &face = mesh->faces[adj_faces[c][i_face]];
      double l[5][5], r[5][5];
      EigenMatrices(big_u, face.frames[0], l, r);
      auto rotated = proj;
      for (auto &pr : rotated) {  // Expansion::LeftMultiply(L)
        for (int i = 0; i < n; ++i) {
          double col[5];
          for (int k = 0; k < 5; ++k) col[k] = pr[5 * i + k];
          for (int k = 0; k < 5; ++k) {
            double v = 0;
            for (int m = 0; m < 5; ++m) v += l[k][m] * col[m];
            pr[5 * i + k] = v;
          }
        }
      }
      Blend(rotated, Weights(c, rotated), tmp.data());
      const int64_t other = face.holder == c ? face.sharer : face.holder;
      const double adj_volume = mesh->cells[other].volume;
      for (int i = 0; i < n; ++i) {
        double col[5];
        for (int k = 0; k < 5; ++k) col[k] = tmp[5 * i + k];
        for (int k = 0; k < 5; ++k) {
          double v = 0;
          for (int m = 0; m < 5; ++m) v += r[k][m] * col[m];
          acc[5 * i + k] += v * adj_volume;
        }
      }
      total_volume += adj_volume;
    }
    for (size_t i = 0; i < Stride(); ++i) out[i] = acc[i] / total_volume;
  }
  // limiter/reconstruct.hpp:22-47: flag all cells, reconstruct the troubled
  // ones from the old state, then write.  Returns the number of troubled cells.
  int64_t Limit(double *u) {
    const size_t nc = mesh->cells.size();
    flags.assign(nc, 0);
    if (limiter == 0) return 0;
    std::vector<int64_t> troubled;
    for (size_t c = 0; c < nc; ++c) {
      flags[c] = (limiter == 2 || limiter == 4) ? 1 : (IsNotSmooth(c, u) ? 1 : 0);
      if (flags[c]) troubled.push_back(c);
    }
    std::vector<double> fresh(troubled.size() * Stride());
    for (size_t t = 0; t < troubled.size(); ++t) {
      if (limiter >= 3) EigenReconstruct(troubled[t], u, fresh.data() + t * Stride());
      else LazyReconstruct(troubled[t], u, fresh.data() + t * Stride());
    }
    for (size_t t = 0; t < troubled.size(); ++t)
      std::copy(fresh.begin() + t * Stride(), fresh.begin() + (t + 1) * Stride(), u + troubled[t] * Stride());
    return (int64_t)troubled.size();
  }

  // fem.hpp:417-435: the first N quadrature points of each cell are sampled.
  double TimeStep(double const *u, double cfl, double dt_guess) const {
    double min_dt = 1e100;
    for (size_t c = 0; c < mesh->cells.size(); ++c) {
      double max_speed = 1e-5;
      for (int i = 0; i < n; ++i) max_speed = std::max(max_speed, GetMaximumSpeed(Value(c, u, mesh->cells[c].qp[i])));
      min_dt = std::min(min_dt, mesh->cells[c].length / max_speed);
    }
    return std::min(cfl * min_dt, dt_guess);
  }
};

}  // namespace oracle

#endif  // ORACLE_DG_HPP_
