// Oracle (test infrastructure): nodal flux reconstruction on hexahedra,
// restating spatial/fr/general.hpp and spatial/fem.hpp.
//
// The reference's fr::General asserts that every face flux point is also a
// solution point (Gauss-Lobatto nodes, fr/general.hpp:107-192).  On
// Gauss-Legendre nodes the flux point lies at the end of the line of
// solution points normal to the face; the natural generalisation used here
// (and by the product) evaluates both the solution and the cached local flux
// F~ = F J* through the line's Lagrange interpolant at that end point.  With
// Lobatto nodes the interpolation weights are exactly (1, 0, ..., 0) and the
// arithmetic reduces to fr/general.hpp:296-315 verbatim.
#ifndef ORACLE_FR_HPP_
#define ORACLE_FR_HPP_

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <vector>

#include "core.hpp"
#include "mesh.hpp"

namespace oracle {

enum BcKind { kBcSmart = 0, kBcInviscidWall = 1, kBcSupersonicOutlet = 2, kBcSupersonicInlet = 3,
              kBcSubsonicInlet = 4, kBcSubsonicOutlet = 5 };
// fem.hpp:261-323: the order in which the boundary families are added.
constexpr int kBcFamilyOrder[] = { kBcSupersonicInlet, kBcSupersonicOutlet, kBcSubsonicInlet, kBcSubsonicOutlet,
                                   kBcInviscidWall, kBcSmart };
struct BcEntry { int kind; Val5 state; };

// Line (Q solution points normal to a face) of local face i_face through
// the flux point with in-face node indices (a, b); returns the Q node ids.
// The in-face index pairs follow fr/general.hpp:103-205:
//   face 0 (zeta=-1): (i, j); 1 (eta=-1): (i, k); 2 (xi=+1): (j, k);
//   face 3 (eta=+1): (i, k); 4 (xi=-1): (j, k); 5 (zeta=+1): (i, j).
inline void FaceLine(int q, int i_face, int a, int b, int *ijk) {
  for (int m = 0; m < q; ++m) {
    switch (i_face) {
    case 0: case 5: ijk[m] = (a * q + b) * q + m; break;
    case 1: case 3: ijk[m] = (a * q + m) * q + b; break;
    default: ijk[m] = (m * q + a) * q + b; break;
    }
  }
}
constexpr int kFaceDir[6] = { 2, 1, 0, 1, 0, 2 };        // xyz normal to the face
constexpr double kFaceSign[6] = { -1, -1, +1, +1, -1, +1 };  // end of the line

struct FrScheme {
  Mesh const *mesh;
  int q, n;  // 1-D points, nodes per cell
  Lagrange1D line;
  Vincent vincent;
  int riemann;
  std::vector<BcEntry> bcs;  // indexed by boundary tag
  // per cell, per node: J* = det(J) J^{-1} and det(J) (fr/general.hpp:249-268)
  std::vector<Mat3> jstar;
  std::vector<double> det;
  // per face point: (holder ijk, sharer ijk) of the node nearest the face,
  // as fr/general.hpp:106-191 stores in FluxPointCache::ijk
  std::vector<std::array<int, 4>> fp_index;  // holder (a, b), sharer (a, b)
  std::vector<RotatedEuler> riemanns;        // [face][point], fem.hpp:95-115
  double dmat[8][8];                         // dmat[i][m] = l_m'(x_i)
  double lend[2][8];                         // l_m(-1), l_m(+1)
  double gprime[2][8];                       // g_L'(x_m), g_R'(x_m)
  // Planted mistakes for the tests that show a pin catches them (0 = none):
  // 1 lifts every flux gap onto the transposed line (b, a); 2 takes the
  // sharer's trace from its transposed line.
  int mutation = 0;
  // Given states per boundary face point, [boundary face][point] (the BC
  // functions f(x, t) evaluated by the caller at the time of the residual);
  // empty: each tag's constant state.
  std::vector<Val5> point_state;

  FrScheme(Mesh const *m, double c_next, int riemann_kind, std::vector<BcEntry> bc)
      : mesh(m), q(m->q), n(m->q * m->q * m->q), vincent(m->q - 1, c_next),
        riemann(riemann_kind), bcs(std::move(bc)) {
    line.q = q;
    for (int i = 0; i < q; ++i) line.x[i] = mesh->x1d[i];
    for (int i = 0; i < q; ++i) for (int m2 = 0; m2 < q; ++m2) dmat[i][m2] = line.Derivative(m2, line.x[i]);
    for (int m2 = 0; m2 < q; ++m2) {
      lend[0][m2] = line.Value(m2, -1.0);
      lend[1][m2] = line.Value(m2, +1.0);
      gprime[0][m2] = vincent.LeftDerivative(line.x[m2]);
      gprime[1][m2] = vincent.RightDerivative(line.x[m2]);
    }
    size_t nc = mesh->cells.size();
    jstar.resize(nc * n);
    det.resize(nc * n);
    for (size_t c = 0; c < nc; ++c) {
      auto const &cell = mesh->cells[c];
      for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) for (int k = 0; k < q; ++k) {
        int ijk = (i * q + j) * q + k;
        Mat3 jac = cell.map.LocalToJacobian(line.x[i], line.x[j], line.x[k]);
        Mat3 inv = inv3(jac);
        double d = det3(jac);
        for (int r = 0; r < 3; ++r) for (int s = 0; s < 3; ++s) jstar[c * n + ijk][r][s] = d * inv[r][s];
        det[c * n + ijk] = cell.qw[ijk] / cell.qw_local[ijk];  // element.hpp:84-86
      }
    }
    int nfp = q * q;
    fp_index.resize(mesh->faces.size() * nfp);
    riemanns.resize(mesh->faces.size() * nfp);
    for (size_t f = 0; f < mesh->faces.size(); ++f) {
      auto const &face = mesh->faces[f];
      for (int p = 0; p < nfp; ++p) {
        riemanns[f * nfp + p].kind = RiemannKind(riemann);
        int a, b;
        FindCollinear(mesh->cells[face.holder], face.holder_face, face.qp[p], &a, &b);
        fp_index[f * nfp + p] = { a, b, -1, -1 };
        if (face.sharer >= 0) {
          FindCollinear(mesh->cells[face.sharer], face.sharer_face, add(face.qp[p], face.shift), &a, &b);
          fp_index[f * nfp + p][2] = a;
          fp_index[f * nfp + p][3] = b;
        }
      }
    }
  }

  // polynomial/hexahedron.hpp:821-928 FindCollinearIndexByGlobal.
  void FindCollinear(CellGeom const &cell, int i_face, Vec3 const &x, int *a, int *b) const {
    double const *p = mesh->x1d;
    for (int s = 0; s < q; ++s) for (int t = 0; t < q; ++t) {
      Vec3 y;
      switch (i_face) {
      case 0: y = cell.map.LocalToGlobal(p[s], p[t], -1); break;
      case 1: y = cell.map.LocalToGlobal(p[s], -1, p[t]); break;
      case 2: y = cell.map.LocalToGlobal(+1, p[s], p[t]); break;
      case 3: y = cell.map.LocalToGlobal(p[s], +1, p[t]); break;
      case 4: y = cell.map.LocalToGlobal(-1, p[s], p[t]); break;
      default: y = cell.map.LocalToGlobal(p[s], p[t], +1); break;
      }
      if (norm(sub(y, x)) < 1e-8) { *a = s; *b = t; return; }
    }
    throw std::runtime_error("FindCollinear: no match");
  }

  // The node nearest the face on the line, i.e. FluxPointCache::ijk.
  int FluxPointNode(int i_face, int a, int b) const {
    int ijk[8];
    FaceLine(q, i_face, a, b, ijk);
    return kFaceSign[i_face] < 0 ? ijk[0] : ijk[q - 1];
  }

  // J* at the flux point: local coordinates of the line's end.
  Mat3 JstarAt(CellGeom const &cell, int i_face, int a, int b) const {
    double const *p = mesh->x1d;
    double e = kFaceSign[i_face];
    double l[3];
    switch (i_face) {
    case 0: case 5: l[0] = p[a]; l[1] = p[b]; l[2] = e; break;
    case 1: case 3: l[0] = p[a]; l[1] = e; l[2] = p[b]; break;
    default: l[0] = e; l[1] = p[a]; l[2] = p[b]; break;
    }
    Mat3 jac = cell.map.LocalToJacobian(l[0], l[1], l[2]);
    Mat3 inv = inv3(jac);
    double d = det3(jac);
    Mat3 r;
    for (int s = 0; s < 3; ++s) for (int t = 0; t < 3; ++t) r[s][t] = d * inv[s][t];
    return r;
  }

  // Column layout: u[(c * n + ijk) * 5 + k] (part.hpp:754-759, expansion.hpp:83-104).
  void Residual(double const *u, double *r) {
    size_t nc = mesh->cells.size();
    std::vector<double> flux(nc * n * 15);  // F~ cache, fr/general.hpp:215
    std::fill(r, r + nc * n * 5, 0.0);
    // AddFluxDivergence, fr/general.hpp:271-288
    for (size_t c = 0; c < nc; ++c) {
      for (int ijk = 0; ijk < n; ++ijk) {
        Val5 uq;
        for (int k = 0; k < 5; ++k) uq[k] = u[(c * n + ijk) * 5 + k];
        Flux53 f = GetFluxMatrix(uq);
        Mat3 const &js = jstar[c * n + ijk];
        for (int k = 0; k < 5; ++k) for (int d = 0; d < 3; ++d)
          flux[((c * n + ijk) * 5 + k) * 3 + d] = f[k][0] * js[0][d] + f[k][1] * js[1][d] + f[k][2] * js[2][d];
      }
      for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) for (int k = 0; k < q; ++k) {
        int ijk = (i * q + j) * q + k;
        for (int v = 0; v < 5; ++v) {
          double val = 0;
          for (int m = 0; m < q; ++m) {
            val += flux[((c * n + (m * q + j) * q + k) * 5 + v) * 3 + 0] * dmat[i][m];
            val += flux[((c * n + (i * q + m) * q + k) * 5 + v) * 3 + 1] * dmat[j][m];
            val += flux[((c * n + (i * q + j) * q + m) * 5 + v) * 3 + 2] * dmat[k][m];
          }
          r[(c * n + ijk) * 5 + v] -= val;
        }
      }
    }
    int nfp = q * q;
    auto trace = [&](int64_t c, int i_face, int a, int b, Val5 *uf, Val5 *fn) {
      int ijk[8];
      FaceLine(q, i_face, a, b, ijk);
      double const *l = lend[kFaceSign[i_face] > 0];
      int dir = kFaceDir[i_face];
      uf->fill(0); fn->fill(0);
      for (int m = 0; m < q; ++m)
        for (int v = 0; v < 5; ++v) {
          (*uf)[v] += l[m] * u[(c * n + ijk[m]) * 5 + v];
          (*fn)[v] += l[m] * flux[((c * n + ijk[m]) * 5 + v) * 3 + dir];
        }
    };
    auto correct = [&](int64_t c, int i_face, int a, int b, Val5 const &gap) {
      int ijk[8];
      if (mutation == 1) std::swap(a, b);
      FaceLine(q, i_face, a, b, ijk);
      double const *g = gprime[kFaceSign[i_face] > 0];
      for (int m = 0; m < q; ++m)
        for (int v = 0; v < 5; ++v) r[(c * n + ijk[m]) * 5 + v] -= gap[v] * g[m];
    };
    auto face_scale = [&](int64_t c, int i_face, int a, int b) {
      Mat3 js = JstarAt(mesh->cells[c], i_face, a, b);
      int dir = kFaceDir[i_face];
      Vec3 col{ js[0][dir], js[1][dir], js[2][dir] };
      return kFaceSign[i_face] * norm(col);
    };
    // AddFluxOnLocalFaces -> AddFluxToHolderAndSharer, fr/general.hpp:409-434
    for (int64_t f = 0; f < mesh->n_interior; ++f) {
      auto const &face = mesh->faces[f];
      for (int p = 0; p < nfp; ++p) {
        auto const &ix = fp_index[f * nfp + p];
        Val5 uh, fh, us, fs;
        trace(face.holder, face.holder_face, ix[0], ix[1], &uh, &fh);
        if (mutation == 2) trace(face.sharer, face.sharer_face, ix[3], ix[2], &us, &fs);
        else trace(face.sharer, face.sharer_face, ix[2], ix[3], &us, &fs);
        Val5 fu = riemanns[f * nfp + p].GetFluxUpwind(face.frames[p], uh, us);
        double sh = face_scale(face.holder, face.holder_face, ix[0], ix[1]);
        double ss = face_scale(face.sharer, face.sharer_face, ix[2], ix[3]);
        Val5 gh, gs;
        for (int v = 0; v < 5; ++v) {
          gh[v] = fu[v] * sh - fh[v];
          gs[v] = fu[v] * (-ss) - fs[v];
        }
        correct(face.holder, face.holder_face, ix[0], ix[1], gh);
        correct(face.sharer, face.sharer_face, ix[2], ix[3], gs);
      }
    }
    // AddFluxOnBoundaries, fem.hpp:260-307 (family order), fr/general.hpp:435-640
    for (int family : kBcFamilyOrder) {
      for (size_t tag = 0; tag < bcs.size(); ++tag) {
        if (bcs[tag].kind != family) continue;
        for (size_t f = mesh->n_interior; f < mesh->faces.size(); ++f) {
          auto const &face = mesh->faces[f];
          if (face.tag != (int)tag) continue;
          for (int p = 0; p < nfp; ++p) {
            auto const &ix = fp_index[f * nfp + p];
            Val5 uh, fh;
            trace(face.holder, face.holder_face, ix[0], ix[1], &uh, &fh);
            auto &rs = riemanns[f * nfp + p];
            Val5 const &given = point_state.empty() ? bcs[tag].state
                                                    : point_state[(f - mesh->n_interior) * nfp + p];
            Val5 fu;
            switch (family) {
            case kBcInviscidWall: fu = rs.GetFluxOnInviscidWall(face.frames[p], uh); break;
            case kBcSupersonicOutlet: fu = rs.GetFluxOfState(face.frames[p], uh); break;
            case kBcSupersonicInlet: fu = rs.GetFluxOfState(face.frames[p], given); break;
            case kBcSubsonicInlet: fu = rs.GetFluxOnSubsonicInlet(face.frames[p], uh, given); break;
            case kBcSubsonicOutlet: fu = rs.GetFluxOnSubsonicOutlet(face.frames[p], uh, given); break;
            default: fu = rs.GetFluxOnSmartBoundary(face.frames[p], uh, given); break;
            }
            double sh = face_scale(face.holder, face.holder_face, ix[0], ix[1]);
            Val5 gh;
            for (int v = 0; v < 5; ++v) gh[v] = fu[v] * sh - fh[v];
            correct(face.holder, face.holder_face, ix[0], ix[1], gh);
          }
        }
      }
    }
    // divide by det(J) at each node, fr/general.hpp:249-268
    for (size_t c = 0; c < nc; ++c)
      for (int ijk = 0; ijk < n; ++ijk) {
        double s = 1.0 / det[c * n + ijk];
        for (int v = 0; v < 5; ++v) r[(c * n + ijk) * 5 + v] *= s;
      }
  }

  // Residual with its loops spread over `threads` host threads, for the CPU
  // baseline: cells split into ranges for the volume term, faces for the
  // fluxes and gaps, then every cell gathers its faces' corrections in the
  // order Residual applies them (interior faces by id; boundary faces by
  // family, tag and id; points in order), so the result is bitwise that of
  // Residual.
  void ResidualThreaded(double const *u, double *r, int threads) {
    if (mutation) throw std::runtime_error("planted mistakes run serially");
    const size_t nc = mesh->cells.size();
    const int nfp = q * q;
    const size_t nf = mesh->faces.size();
    auto parallel = [threads](size_t count, auto &&body) {
      const int t_n = std::max(1, std::min<int>(threads, (int)count));
      std::vector<std::thread> pool;
      for (int t = 0; t < t_n; ++t)
        pool.emplace_back([&, t] { body(count * t / t_n, count * (t + 1) / t_n); });
      for (auto &th : pool) th.join();
    };
    if (cell_faces_.empty()) {  // each cell's (face, holder?) pairs in the order Residual visits them
      cell_faces_.assign(nc, {});
      for (int64_t f = 0; f < mesh->n_interior; ++f) {
        cell_faces_[mesh->faces[f].holder].push_back({ f, true });
        cell_faces_[mesh->faces[f].sharer].push_back({ f, false });
      }
      for (int family : kBcFamilyOrder)
        for (size_t tag = 0; tag < bcs.size(); ++tag) {
          if (bcs[tag].kind != family) continue;
          for (size_t f = mesh->n_interior; f < nf; ++f)
            if (mesh->faces[f].tag == (int)tag) cell_faces_[mesh->faces[f].holder].push_back({ (int64_t)f, true });
        }
    }
    if (!flux_) {  // reused, every entry is written before it is read
      flux_ = std::make_unique_for_overwrite<double[]>(nc * n * 15);
      gaps_ = std::make_unique_for_overwrite<Val5[]>(nf * nfp * 2);
    }
    double *flux = flux_.get();
    Val5 *gaps = gaps_.get();  // [face][point][holder, sharer]
    parallel(nc, [&](size_t c0, size_t c1) {  // AddFluxDivergence, fr/general.hpp:271-288
      std::fill(r + c0 * n * 5, r + c1 * n * 5, 0.0);
      for (size_t c = c0; c < c1; ++c) {
        for (int ijk = 0; ijk < n; ++ijk) {
          Val5 uq;
          for (int k = 0; k < 5; ++k) uq[k] = u[(c * n + ijk) * 5 + k];
          Flux53 f = GetFluxMatrix(uq);
          Mat3 const &js = jstar[c * n + ijk];
          for (int k = 0; k < 5; ++k) for (int d = 0; d < 3; ++d)
            flux[((c * n + ijk) * 5 + k) * 3 + d] = f[k][0] * js[0][d] + f[k][1] * js[1][d] + f[k][2] * js[2][d];
        }
        for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) for (int k = 0; k < q; ++k) {
          int ijk = (i * q + j) * q + k;
          for (int v = 0; v < 5; ++v) {
            double val = 0;
            for (int m = 0; m < q; ++m) {
              val += flux[((c * n + (m * q + j) * q + k) * 5 + v) * 3 + 0] * dmat[i][m];
              val += flux[((c * n + (i * q + m) * q + k) * 5 + v) * 3 + 1] * dmat[j][m];
              val += flux[((c * n + (i * q + j) * q + m) * 5 + v) * 3 + 2] * dmat[k][m];
            }
            r[(c * n + ijk) * 5 + v] -= val;
          }
        }
      }
    });
    auto trace = [&](int64_t c, int i_face, int a, int b, Val5 *uf, Val5 *fn) {
      int ijk[8];
      FaceLine(q, i_face, a, b, ijk);
      double const *l = lend[kFaceSign[i_face] > 0];
      int dir = kFaceDir[i_face];
      uf->fill(0); fn->fill(0);
      for (int m = 0; m < q; ++m)
        for (int v = 0; v < 5; ++v) {
          (*uf)[v] += l[m] * u[(c * n + ijk[m]) * 5 + v];
          (*fn)[v] += l[m] * flux[((c * n + ijk[m]) * 5 + v) * 3 + dir];
        }
    };
    auto face_scale = [&](int64_t c, int i_face, int a, int b) {
      Mat3 js = JstarAt(mesh->cells[c], i_face, a, b);
      int dir = kFaceDir[i_face];
      Vec3 col{ js[0][dir], js[1][dir], js[2][dir] };
      return kFaceSign[i_face] * norm(col);
    };
    parallel(nf, [&](size_t f0, size_t f1) {
      for (size_t f = f0; f < f1; ++f) {
        auto const &face = mesh->faces[f];
        for (int p = 0; p < nfp; ++p) {
          auto const &ix = fp_index[f * nfp + p];
          Val5 uh, fh;
          trace(face.holder, face.holder_face, ix[0], ix[1], &uh, &fh);
          Val5 fu;
          if ((int64_t)f < mesh->n_interior) {
            Val5 us, fs;
            trace(face.sharer, face.sharer_face, ix[2], ix[3], &us, &fs);
            fu = riemanns[f * nfp + p].GetFluxUpwind(face.frames[p], uh, us);
            double ss = face_scale(face.sharer, face.sharer_face, ix[2], ix[3]);
            for (int v = 0; v < 5; ++v) gaps[(f * nfp + p) * 2 + 1][v] = fu[v] * (-ss) - fs[v];
          } else {
            auto &rs = riemanns[f * nfp + p];
            const int family = bcs[face.tag].kind;
            Val5 const &given = point_state.empty() ? bcs[face.tag].state
                                                    : point_state[(f - mesh->n_interior) * nfp + p];
            switch (family) {
            case kBcInviscidWall: fu = rs.GetFluxOnInviscidWall(face.frames[p], uh); break;
            case kBcSupersonicOutlet: fu = rs.GetFluxOfState(face.frames[p], uh); break;
            case kBcSupersonicInlet: fu = rs.GetFluxOfState(face.frames[p], given); break;
            case kBcSubsonicInlet: fu = rs.GetFluxOnSubsonicInlet(face.frames[p], uh, given); break;
            case kBcSubsonicOutlet: fu = rs.GetFluxOnSubsonicOutlet(face.frames[p], uh, given); break;
            default: fu = rs.GetFluxOnSmartBoundary(face.frames[p], uh, given); break;
            }
          }
          double sh = face_scale(face.holder, face.holder_face, ix[0], ix[1]);
          for (int v = 0; v < 5; ++v) gaps[(f * nfp + p) * 2][v] = fu[v] * sh - fh[v];
        }
      }
    });
    parallel(nc, [&](size_t c0, size_t c1) {
      for (size_t c = c0; c < c1; ++c) {
        for (auto const &[f, holder] : cell_faces_[c]) {
          auto const &face = mesh->faces[f];
          const int i_face = holder ? face.holder_face : face.sharer_face;
          double const *g = gprime[kFaceSign[i_face] > 0];
          for (int p = 0; p < nfp; ++p) {
            auto const &ix = fp_index[f * nfp + p];
            const int a = holder ? ix[0] : ix[2], b = holder ? ix[1] : ix[3];
            Val5 const &gap = gaps[(f * nfp + p) * 2 + (holder ? 0 : 1)];
            int ijk[8];
            FaceLine(q, i_face, a, b, ijk);
            for (int m = 0; m < q; ++m)
              for (int v = 0; v < 5; ++v) r[(c * n + ijk[m]) * 5 + v] -= gap[v] * g[m];
          }
        }
        for (int ijk = 0; ijk < n; ++ijk) {  // divide by det(J) at each node, fr/general.hpp:249-268
          double sc = 1.0 / det[c * n + ijk];
          for (int v = 0; v < 5; ++v) r[(c * n + ijk) * 5 + v] *= sc;
        }
      }
    });
  }
  // filled by ResidualThreaded
  std::vector<std::vector<std::pair<int64_t, bool>>> cell_faces_;
  std::unique_ptr<double[]> flux_;
  std::unique_ptr<Val5[]> gaps_;

  // limiter/average.hpp:11-28 (kind 0): every node takes the cell average
  // sum_q u_q w_q / volume (polynomial/hexahedron.hpp:626-632);
  // limiter/majority.hpp:14-42 (kind 1, degree >= 2): every node takes, per
  // component, the first value whose running count exceeds N / 2.
  void StartupLimiter(double *u, int kind) const {
    if (kind == 1 && q - 1 < 2) return;
    for (size_t c = 0; c < mesh->cells.size(); ++c) {
      auto const &cell = mesh->cells[c];
      double *uc = u + c * n * 5;
      Val5 value{};
      if (kind == 0) {
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < 5; ++k) value[k] += uc[i * 5 + k] * cell.qw[i];
        for (int k = 0; k < 5; ++k) value[k] /= cell.volume;
      } else {
        for (int k = 0; k < 5; ++k) {
          std::unordered_map<double, int> count;
          bool found = false;
          for (int i = 0; i < n && !found; ++i) {
            double v = uc[i * 5 + k];
            if (++count[v] * 2 > n) { value[k] = v; found = true; }
          }
          if (!found) throw std::runtime_error("majority: no value held by more than half of the nodes");
        }
      }
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < 5; ++k) uc[i * 5 + k] = value[k];
    }
  }

  // fem.hpp:417-435 with an explicit CFL number: every node is sampled.
  double TimeStep(double const *u, double cfl, double dt_guess) const {
    double min_dt = 1e100;
    for (size_t c = 0; c < mesh->cells.size(); ++c) {
      double max_speed = 1e-5;
      for (int i = 0; i < n; ++i) {
        Val5 uq;
        for (int k = 0; k < 5; ++k) uq[k] = u[(c * n + i) * 5 + k];
        max_speed = std::max(max_speed, GetMaximumSpeed(uq));
      }
      min_dt = std::min(min_dt, mesh->cells[c].length / max_speed);
    }
    return std::min(cfl * min_dt, dt_guess);
  }
};

}  // namespace oracle

#endif  // ORACLE_FR_HPP_
