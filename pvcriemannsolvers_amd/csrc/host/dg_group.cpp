// DgGroup: host construction and device upload of one modal-DG mesh.
#include "dg_group.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>

#include "hex_connectivity.hpp"
#include "simplex_rules.hpp"
#include "tables.hpp"

namespace mcfd {

namespace {

using Vec = std::array<double, 3>;

// CGNS corner coordinates of the reference elements and their face tables
// (coordinate/hexahedron.hpp:353-389, coordinate/tetrahedron.hpp:136-148).
constexpr double kHexLocal[8][3] = {
  { -1, -1, -1 }, { +1, -1, -1 }, { +1, +1, -1 }, { -1, +1, -1 },
  { -1, -1, +1 }, { +1, -1, +1 }, { +1, +1, +1 }, { -1, +1, +1 },
};
constexpr int kHexFaces[6][4] = {
  { 0, 3, 2, 1 }, { 0, 1, 5, 4 }, { 1, 2, 6, 5 }, { 2, 3, 7, 6 }, { 0, 4, 7, 3 }, { 4, 5, 6, 7 },
};
constexpr double kQuadLocal[4][2] = { { -1, -1 }, { +1, -1 }, { +1, +1 }, { -1, +1 } };
constexpr double kTetLocal[4][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 }, { 0, 0, 0 } };
constexpr int kTetFaces[4][3] = { { 0, 2, 1 }, { 0, 1, 3 }, { 0, 3, 2 }, { 1, 2, 3 } };

// Position of holder-sorted face node i in the other cell's sorted face list:
// rotations, then reflections, of an n-gon.
int Perm(int n, int orient, int i) {
  const int r = orient % n;
  return orient < n ? (i + r) % n : ((r - i) % n + n) % n;
}

Vec Sub(Vec a, const Vec &b) { for (int d = 0; d < 3; ++d) a[d] -= b[d]; return a; }
double Dot(const Vec &a, const Vec &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec Cross(const Vec &a, const Vec &b) {
  return { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
}
Vec Normalized(const Vec &a) {
  const double s = 1.0 / std::sqrt(Dot(a, a));
  return { a[0] * s, a[1] * s, a[2] * s };
}
double Det3(const double j[9]) {
  return j[0] * (j[4] * j[8] - j[5] * j[7]) - j[1] * (j[3] * j[8] - j[5] * j[6]) + j[2] * (j[3] * j[7] - j[4] * j[6]);
}

// Faces of a conforming tetrahedral mesh in the reference's order: interior
// faces by (holder, sharer, holder face) with holder = the lower cell id
// (mesh/part.hpp:932-975), then boundary faces in input order (:1456-1549).
std::vector<DgFace> BuildTetFaces(const mcfd_mesh_desc &mesh, int64_t *n_interior) {
  using Key = std::array<int64_t, 3>;
  auto key_of = [](int64_t a, int64_t b, int64_t c) {
    Key k{ a, b, c };
    std::sort(k.begin(), k.end());
    return k;
  };
  std::map<Key, std::vector<std::pair<int64_t, int>>> by_key;
  for (int64_t c = 0; c < mesh.n_cells; ++c) {
    const int64_t *cn = mesh.cell_nodes + 4 * c;
    for (int f = 0; f < 4; ++f)
      by_key[key_of(cn[kTetFaces[f][0]], cn[kTetFaces[f][1]], cn[kTetFaces[f][2]])].push_back({ c, f });
  }
  std::vector<std::tuple<int64_t, int64_t, int, int>> inner;
  std::map<Key, std::pair<int64_t, int>> open;
  for (auto &[key, list] : by_key) {
    if (list.size() == 2) {
      auto a = list[0], b = list[1];
      if (b.first < a.first) std::swap(a, b);
      if (a.first == b.first) throw std::invalid_argument("a tetrahedron shares a face with itself");
      inner.emplace_back(a.first, b.first, a.second, b.second);
    } else if (list.size() == 1) {
      open[key] = list[0];
    } else {
      throw std::invalid_argument("a triangle is shared by more than two tetrahedra");
    }
  }
  std::sort(inner.begin(), inner.end());
  std::vector<DgFace> faces;
  for (auto &[h, s, fh, fs] : inner) faces.push_back(DgFace{ h, s, fh, fs, 0, -1 });
  *n_interior = (int64_t)faces.size();
  for (int64_t b = 0; b < mesh.n_bfaces; ++b) {
    const int64_t *bn = mesh.bface_nodes + 3 * b;
    if (mesh.bface_tags[b] < 0) throw std::invalid_argument("periodic faces are supported on hexahedra only");
    auto it = open.find(key_of(bn[0], bn[1], bn[2]));
    if (it == open.end()) throw std::invalid_argument("boundary triangle is not an open cell face");
    faces.push_back(DgFace{ it->second.first, -1, it->second.second, -1, 0, mesh.bface_tags[b] });
    open.erase(it);
  }
  if (!open.empty()) throw std::invalid_argument("open cell faces without a boundary entry");
  return faces;
}

// Gram-Schmidt factor S of the Taylor monomials under the cell quadrature
// (integrator/function.hpp:139-168), lower triangular, row-major n x n.
void OrthoNormalize(int p, int n, const std::vector<Vec> &y, const std::vector<double> &w, double *s) {
  double a[10][10] = {};
  for (size_t q = 0; q < y.size(); ++q) {
    double t[10] = { 1, y[q][0], y[q][1], y[q][2] };
    if (p >= 2) {
      t[4] = y[q][0] * y[q][0]; t[5] = y[q][0] * y[q][1]; t[6] = y[q][0] * y[q][2];
      t[7] = y[q][1] * y[q][1]; t[8] = y[q][1] * y[q][2]; t[9] = y[q][2] * y[q][2];
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) a[i][j] += t[i] * t[j] * w[q];
  }
  double m[10][10];
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) m[i][j] = (i == j);
  m[0][0] = 1 / std::sqrt(a[0][0]);
  for (int i = 1; i < n; ++i) {
    for (int j = 0; j < i; ++j) {
      double dotp = 0;
      for (int k = 0; k <= j; ++k) dotp += m[j][k] * a[k][i];
      for (int l = 0; l <= j; ++l) m[i][l] -= dotp * m[j][l];
    }
    double norm_sq = 0;
    for (int j = 0; j <= i; ++j) {
      double cross = 0;
      for (int k = 0; k < j; ++k) cross += 2 * m[i][k] * a[k][j];
      norm_sq += m[i][j] * (m[i][j] * a[j][j] + cross);
    }
    const double len = std::sqrt(norm_sq);
    for (int j = 0; j <= i; ++j) m[i][j] /= len;
  }
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) s[i * n + j] = m[i][j];
}

}  // namespace

void DgGroup::Build(const mcfd_mesh_desc &mesh, const mcfd_scheme_desc &scheme,
                    const std::vector<int32_t> &bc_kinds, const std::vector<double> &bc_states) {
  if (scheme.degree < 1 || scheme.degree > 2) throw std::invalid_argument("modal DG degree must be 1 or 2");
  if (scheme.riemann < 0 || scheme.riemann > 2) throw std::invalid_argument("unknown Riemann solver");
  if (scheme.precision != MCFD_PRECISION_FP64) throw std::invalid_argument("modal DG runs in FP64 only");
  if (mesh.cell_type != MCFD_HEXA_8 && mesh.cell_type != MCFD_TETRA_4) throw std::invalid_argument("unknown cell type");
  cell_type = mesh.cell_type;
  degree = scheme.degree;
  riemann = scheme.riemann;
  if (scheme.limiter < MCFD_LIMITER_NONE || scheme.limiter > MCFD_LIMITER_WENO_EIGEN_ALL)
    throw std::invalid_argument("unknown limiter");
  limiter = scheme.limiter;
  weno_w0 = scheme.weno_w0;
  weno_eps = scheme.weno_eps;
  if (limiter && !(weno_eps > 0)) throw std::invalid_argument("weno_eps must be positive");
  const bool hex = cell_type == MCFD_HEXA_8;
  const int npe = hex ? 8 : 4;
  n = degree == 1 ? 4 : 10;
  nq = hex ? (degree + 1) * (degree + 1) * (degree + 1) : 14;
  nf = hex ? 6 : 4;
  nfp = hex ? (degree + 1) * (degree + 1) : 6;
  n_cells = mesh.n_cells;
  if (n_cells < 1) throw std::invalid_argument("empty mesh");
  if (n_cells > INT32_MAX) throw std::invalid_argument("too many cells for int32 neighbour ids");
  const int n_bcs = (int)bc_kinds.size();
  for (int64_t b = 0; b < mesh.n_bfaces; ++b)
    if (mesh.bface_tags[b] >= n_bcs) throw std::invalid_argument("boundary tag without a BC entry");
  for (auto k : bc_kinds)
    if (k < 0 || k > 5) throw std::invalid_argument("unknown BC kind");
  h_bc_kind = bc_kinds;
  h_bc_state = bc_states;
  if (h_bc_kind.empty()) {
    h_bc_kind.push_back(1);
    h_bc_state.assign(5, 0.0);
  }
  auto node = [&](int64_t c, int i) {
    const int64_t id = mesh.cell_nodes[npe * c + i];
    if (id < 0 || id >= mesh.n_nodes) throw std::invalid_argument("cell node id out of range");
    return Vec{ mesh.coords[3 * id], mesh.coords[3 * id + 1], mesh.coords[3 * id + 2] };
  };

  // ---- faces
  std::vector<std::array<double, 3>> shift;
  if (hex) {
    HexConnectivity conn = BuildHexConnectivity(mesh);
    for (auto a : conn.affine)
      if (!a) throw Failure{ MCFD_ENOTSUP, "modal DG supports affine (parallelepiped) hexahedra only; "
                                           "trilinear hexahedra run with FR_GENERAL" };
    n_interior = conn.n_interior;
    for (auto const &f : conn.faces) {
      faces.push_back(DgFace{ f.holder, f.sharer, f.holder_face, f.sharer_face, 0, f.tag });
      shift.push_back({ f.shift[0], f.shift[1], f.shift[2] });
    }
  } else {
    faces = BuildTetFaces(mesh, &n_interior);
    shift.assign(faces.size(), { 0, 0, 0 });
  }
  if ((int64_t)faces.size() >= (int64_t(1) << 23)) throw std::invalid_argument("too many faces for the face id field");
  const int fnpe = hex ? 4 : 3;
  auto sorted_face = [&](int64_t c, int lf, int i) {
    return hex ? node(c, kHexFaces[lf][i]) : node(c, kTetFaces[lf][i]);
  };
  auto sorted_face_id = [&](int64_t c, int lf, int i) {
    return mesh.cell_nodes[npe * c + (hex ? kHexFaces[lf][i] : kTetFaces[lf][i])];
  };
  for (int64_t f = 0; f < n_interior; ++f) {
    auto &face = faces[f];
    const bool periodic = shift[f][0] != 0 || shift[f][1] != 0 || shift[f][2] != 0;
    int found = -1;
    for (int o = 0; o < 2 * fnpe && found < 0; ++o) {
      bool ok = true;
      for (int i = 0; i < fnpe && ok; ++i) {
        const int j = Perm(fnpe, o, i);
        if (periodic) {
          const Vec xh = sorted_face(face.holder, face.holder_face, i);
          const Vec xs = sorted_face(face.sharer, face.sharer_face, j);
          double d2 = 0, l2 = 0;
          for (int d = 0; d < 3; ++d) {
            const double e = xh[d] + shift[f][d] - xs[d];
            d2 += e * e;
            l2 += xh[d] * xh[d];
          }
          ok = d2 <= 1e-20 * (1 + l2);
        } else {
          ok = sorted_face_id(face.holder, face.holder_face, i) == sorted_face_id(face.sharer, face.sharer_face, j);
        }
      }
      if (ok) found = o;
    }
    if (found < 0) throw std::invalid_argument("interior face whose corners do not match");
    face.orient = found;
  }

  // ---- reference tables
  tables = DgTables{};
  double xc[3] = { 0, 0, 0 };
  double x1d[8], w1d[8];
  if (hex) {
    const int q1 = degree + 1;
    GaussRule(false, q1, x1d, w1d);
    for (int i = 0, q = 0; i < q1; ++i)
      for (int j = 0; j < q1; ++j)
        for (int k = 0; k < q1; ++k, ++q) {
          tables.vol_eta[q][0] = x1d[i]; tables.vol_eta[q][1] = x1d[j]; tables.vol_eta[q][2] = x1d[k];
          tables.vol_w[q] = w1d[i] * w1d[j] * w1d[k];
        }
    for (int i = 0, q = 0; i < q1; ++i)
      for (int j = 0; j < q1; ++j, ++q) {
        tables.face_w[q] = w1d[i] * w1d[j];
        for (int lf = 0; lf < 6; ++lf)
          for (int o = 0; o < 8; ++o)
            for (int d = 0; d < 3; ++d) {
              double v = 0;
              for (int c = 0; c < 4; ++c) {
                const double sh = (1 + kQuadLocal[c][0] * x1d[i]) * (1 + kQuadLocal[c][1] * x1d[j]) / 4;
                v += sh * kHexLocal[kHexFaces[lf][Perm(4, o, c)]][d];
              }
              tables.face_eta[lf][o][q][d] = v;
            }
      }
  } else {
    xc[0] = xc[1] = xc[2] = 0.25;
    double tx[14][3], tw[14], fx[6][2], fw[6];
    TetRule14(tx, tw);
    TriRule6(fx, fw);
    for (int q = 0; q < 14; ++q) {
      for (int d = 0; d < 3; ++d) tables.vol_eta[q][d] = tx[q][d] - xc[d];
      tables.vol_w[q] = tw[q];
    }
    for (int q = 0; q < 6; ++q) {
      tables.face_w[q] = fw[q];
      const double sh[3] = { fx[q][0], fx[q][1], 1.0 - fx[q][0] - fx[q][1] };
      for (int lf = 0; lf < 4; ++lf)
        for (int o = 0; o < 6; ++o)
          for (int d = 0; d < 3; ++d) {
            double v = 0;
            for (int c = 0; c < 3; ++c) v += sh[c] * kTetLocal[kTetFaces[lf][Perm(3, o, c)]][d];
            tables.face_eta[lf][o][q][d] = v - xc[d];
          }
    }
  }

  // ---- cells: geometry, quadrature, orthonormal basis
  const int packed = n * (n + 1) / 2;
  h_geo.assign(n_cells * kDgGeoStride, 0.0);
  h_basis.assign(n_cells * packed, 0.0);
  qp.assign(n_cells * nq * 3, 0.0);
  std::vector<double> s(n * n);
  for (int64_t c = 0; c < n_cells; ++c) {
    Vec x[8];
    for (int i = 0; i < npe; ++i) x[i] = node(c, i);
    double j[9] = {};
    Vec center{ 0, 0, 0 };
    if (hex) {  // Hexahedron8 at the centre: J rows = sum_i xi_i x_i / 8, centre = sum_i x_i / 8
      for (int i = 0; i < 8; ++i)
        for (int d = 0; d < 3; ++d) {
          center[d] += x[i][d] / 8;
          for (int r = 0; r < 3; ++r) j[3 * r + d] += kHexLocal[i][r] * x[i][d] / 8;
        }
    } else {  // Tetrahedron4: rows x_r - x_3, centre at xi = (1/4, 1/4, 1/4)
      for (int r = 0; r < 3; ++r)
        for (int d = 0; d < 3; ++d) j[3 * r + d] = x[r][d] - x[3][d];
      for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) center[d] += x[i][d] * 0.25;
    }
    const double det = std::abs(Det3(j));
    if (!(det > 0)) throw std::invalid_argument("degenerate cell");
    std::vector<Vec> y(nq);
    std::vector<double> w(nq);
    double vol = 0;
    for (int q = 0; q < nq; ++q) {
      for (int d = 0; d < 3; ++d) {
        const double *e = tables.vol_eta[q];
        y[q][d] = e[0] * j[d] + e[1] * j[3 + d] + e[2] * j[6 + d];
        qp[(c * nq + q) * 3 + d] = center[d] + y[q][d];
      }
      w[q] = tables.vol_w[q] * det;
      vol += w[q];
    }
    OrthoNormalize(degree, n, y, w, s.data());
    for (int i = 0, k = 0; i < n; ++i)
      for (int jj = 0; jj <= i; ++jj, ++k) h_basis[c * packed + k] = s[i * n + jj];
    double *g = h_geo.data() + c * kDgGeoStride;
    for (int k = 0; k < 9; ++k) g[k] = j[k];
    g[9] = det;
    g[10] = std::cbrt(vol) * 0.5;  // mesh/part.hpp:230
    for (int d = 0; d < 3; ++d) g[11 + d] = center[d];
    g[14] = vol;
  }

  // ---- face records and per-cell face info
  h_faces.assign(faces.size() * kDgFaceStride, 0.0);
  h_face_cells.assign(faces.size() * 3, 0);
  h_nbr.assign(n_cells * 6, INT32_MIN);
  h_finfo.assign(n_cells * 6, 0);
  for (size_t f = 0; f < faces.size(); ++f) {
    const auto &face = faces[f];
    Vec ra, rb;
    if (hex) {  // Quadrangle4 Jacobian at the face centre
      ra = { 0, 0, 0 }; rb = { 0, 0, 0 };
      for (int c = 0; c < 4; ++c) {
        const Vec xc4 = sorted_face(face.holder, face.holder_face, c);
        for (int d = 0; d < 3; ++d) {
          ra[d] += kQuadLocal[c][0] / 4 * xc4[d];
          rb[d] += kQuadLocal[c][1] / 4 * xc4[d];
        }
      }
    } else {  // Triangle3: rows x_0 - x_2, x_1 - x_2
      const Vec x0 = sorted_face(face.holder, face.holder_face, 0), x1 = sorted_face(face.holder, face.holder_face, 1),
                x2 = sorted_face(face.holder, face.holder_face, 2);
      ra = Sub(x0, x2);
      rb = Sub(x1, x2);
    }
    const Vec nrm = Normalized(Cross(ra, rb)), tan = Normalized(ra);
    double *fr = h_faces.data() + f * kDgFaceStride;
    for (int d = 0; d < 3; ++d) { fr[d] = nrm[d]; fr[3 + d] = tan[d]; }
    const double g11 = Dot(ra, ra), g12 = Dot(ra, rb), g22 = Dot(rb, rb);
    fr[6] = std::sqrt(g11 * g22 - g12 * g12);  // integrator/element.hpp:108-110
    const int32_t fid = (int32_t)f << 8;
    h_face_cells[f * 3] = (int32_t)face.holder;
    h_face_cells[f * 3 + 1] = face.sharer >= 0 ? (int32_t)face.sharer : -1 - face.tag;
    h_face_cells[f * 3 + 2] = face.holder_face | (face.sharer >= 0 ? face.sharer_face : 0) << 3 | face.orient << 6;
    if (face.sharer >= 0) {
      h_nbr[face.holder * 6 + face.holder_face] = (int32_t)face.sharer;
      h_finfo[face.holder * 6 + face.holder_face] = fid | 1 << 7 | face.orient << 3 | face.sharer_face;
      h_nbr[face.sharer * 6 + face.sharer_face] = (int32_t)face.holder;
      h_finfo[face.sharer * 6 + face.sharer_face] = fid | face.orient << 3 | face.holder_face;
    } else {
      h_nbr[face.holder * 6 + face.holder_face] = -1 - face.tag;
      h_finfo[face.holder * 6 + face.holder_face] = fid | 1 << 7;
    }
  }
  for (int64_t c = 0; c < n_cells; ++c)
    for (int lf = 0; lf < nf; ++lf)
      if (h_nbr[c * 6 + lf] == INT32_MIN) throw std::invalid_argument("cell face without a neighbour or boundary");

  // ---- limiter neighbours (mesh/part.hpp:962-963): interior faces in order
  h_adj.assign(n_cells * 6, -1);
  h_adj_face.assign(n_cells * 6, -1);
  h_nadj.assign(n_cells, 0);
  for (int64_t f = 0; f < n_interior; ++f) {
    const auto &face = faces[f];
    if (limiter && (shift[f][0] != 0 || shift[f][1] != 0 || shift[f][2] != 0))
      throw std::invalid_argument("the WENO limiter is not defined across periodic faces");
    h_adj_face[face.holder * 6 + h_nadj[face.holder]] = (int32_t)f;
    h_adj[face.holder * 6 + h_nadj[face.holder]++] = (int32_t)face.sharer;
    h_adj_face[face.sharer * 6 + h_nadj[face.sharer]] = (int32_t)f;
    h_adj[face.sharer * 6 + h_nadj[face.sharer]++] = (int32_t)face.holder;
  }
}

void DgGroup::Upload(hipStream_t stream) {
  const size_t cs = ColumnSize();
  u0.Allocate(cs); u1.Allocate(cs); u2.Allocate(cs); res.Allocate(cs);
  ct0.Allocate(cs); ct1.Allocate(cs); ct2.Allocate(cs);
  Check(hipMemsetAsync(ct0.ptr, 0, cs * sizeof(double), stream), "memset");
  Check(hipMemsetAsync(u0.ptr, 0, cs * sizeof(double), stream), "memset");
  basis.Upload(h_basis, stream);
  geo.Upload(h_geo, stream);
  face_rec.Upload(h_faces, stream);
  face_cells.Upload(h_face_cells, stream);
  fhat.Allocate(faces.size() * nfp * 5);
  if (riemann == MCFD_RIEMANN_EXACT) {
    exact_star.Allocate(faces.size() * nfp);
    Check(hipMemsetAsync(exact_star.ptr, 0, faces.size() * nfp, stream), "memset");
  }
  bc_state.Upload(h_bc_state, stream);
  finfo.Upload(h_finfo, stream);
  bc_kind.Upload(h_bc_kind, stream);
  h_tables.assign(1, tables);
  tables_d.Upload(h_tables, stream);
  if (limiter) {
    adj.Upload(h_adj, stream);
    adj_face.Upload(h_adj_face, stream);
    nadj.Upload(h_nadj, stream);
    flags.Allocate(n_cells);
    Check(hipMemsetAsync(flags.ptr, 0, n_cells, stream), "memset");
    std::vector<double> h_rec(size_t(n_cells) * kDgFlagRecord, 0.0);  // geometry terms; the rest per state
    for (int64_t c = 0; c < n_cells; ++c) {
      for (int d = 0; d < 3; ++d) h_rec[c * kDgFlagRecord + kDgFlagCentre + d] = h_geo[c * kDgGeoStride + 11 + d];
      h_rec[c * kDgFlagRecord + kDgFlagVolume] = h_geo[c * kDgGeoStride + 14];
      h_rec[c * kDgFlagRecord + kDgFlagDivisor] = std::pow(h_geo[c * kDgGeoStride + 14], (degree + 1.0) / 2.0 / 3);
      const double s00 = h_basis[c * (n * (n + 1) / 2)];
      const double measure = 1 / (s00 * s00);  // OrthoNormal::Measure
      h_rec[c * kDgFlagRecord + kDgFlagScale1] = std::pow(measure, 2. / 3 - 1);
      h_rec[c * kDgFlagRecord + kDgFlagScale2] = std::pow(measure, 4. / 3 - 1);
    }
    flag_rec.Upload(h_rec, stream);
    troubled.Allocate(size_t(n_cells) + 1);
    proj.Allocate(size_t(n_cells) * 7 * 5 * n);
    if (limiter >= MCFD_LIMITER_WENO_EIGEN) eig.Allocate(size_t(n_cells) * 6 * 5 * n);
  }
}

void DgGroup::Limit(double *u, double *ct, hipStream_t stream) {
  if (!limiter) return;
  DgLimiterArgs a;
  a.u = u; a.ct = ct; a.basis = basis.ptr; a.geo = geo.ptr;
  a.adj = adj.ptr; a.adj_face = adj_face.ptr; a.n_adj = nadj.ptr; a.faces = face_rec.ptr; a.tables = tables_d.ptr;
  a.flag_rec = flag_rec.ptr; a.flags = flags.ptr; a.troubled = troubled.ptr; a.n_troubled = troubled.ptr + n_cells;
  a.proj = proj.ptr; a.eig = eig.ptr;
  a.n_cells = n_cells; a.mode = limiter; a.w0 = weno_w0; a.eps = weno_eps;
  Check(LaunchDgLimiter(cell_type, degree, a, stream), "DG limiter launch");
}

void DgGroup::SolutionChanged(hipStream_t stream, bool limit) {
  Check(LaunchDgTaylorCoefficients(degree, u0.ptr, basis.ptr, n_cells, ct0.ptr, limiter ? flag_rec.ptr : nullptr,
                                   stream),
        "DG Taylor launch");
  if (limit) Limit(u0.ptr, ct0.ptr, stream);  // spatial/with_limiter.hpp:56-59
}

void DgGroup::Stage(int stage, double dt, hipStream_t stream, unsigned long long *dt_min_bits) {
  DgModalArgs a;
  a.u0 = u0.ptr;
  switch (stage) {
  case 0: a.us = u0.ptr; a.ct = ct0.ptr; a.out = res.ptr; a.ct_out = nullptr; break;
  case 1: a.us = u0.ptr; a.ct = ct0.ptr; a.out = u1.ptr; a.ct_out = ct1.ptr; break;
  case 2: a.us = u1.ptr; a.ct = ct1.ptr; a.out = u2.ptr; a.ct_out = ct2.ptr; break;
  default: a.us = u2.ptr; a.ct = ct2.ptr; a.out = u0.ptr; a.ct_out = ct0.ptr; break;
  }
  a.basis = basis.ptr; a.geo = geo.ptr;
  a.finfo = finfo.ptr; a.faces = face_rec.ptr;
  a.bc_kind = bc_kind.ptr; a.bc_state = bc_state.ptr; a.tables = tables_d.ptr;
  a.n_cells = n_cells; a.dt = dt; a.stage = stage;
  a.dt_min_bits = stage > 0 && !limiter ? dt_min_bits : nullptr;
  a.fhat = fhat.ptr;
  a.exact_star = exact_star.ptr;
  a.bc_point_state = bc_point_state;
  a.n_interior = n_interior;
  a.flag_rec = limiter && stage > 0 ? flag_rec.ptr : nullptr;
  a.face_cells = face_cells.ptr;
  a.n_faces = (int64_t)faces.size();
  Check(LaunchDgModalStage(cell_type, degree, riemann, a, stream), "DG stage launch");
  if (stage > 0) Limit(a.out, a.ct_out, stream);  // spatial/with_limiter.hpp:56-59 after each SetSolutionColumn
}

void DgGroup::BoundaryPoints(std::vector<int32_t> *tags, std::vector<double> *xyz) const {
  const int64_t nb = (int64_t)faces.size() - n_interior;
  tags->assign(nb, 0);
  xyz->assign(nb * nfp * 3, 0.0);
  for (int64_t b = 0; b < nb; ++b) {
    const auto &face = faces[n_interior + b];
    (*tags)[b] = face.tag;
    const double *g = h_geo.data() + face.holder * kDgGeoStride;  // J rows, then the centre at [11..13]
    for (int q = 0; q < nfp; ++q) {
      const double *e = tables.face_eta[face.holder_face][0][q];
      for (int d = 0; d < 3; ++d)
        (*xyz)[(b * nfp + q) * 3 + d] = g[11 + d] + (e[0] * g[d] + e[1] * g[3 + d] + e[2] * g[6 + d]);
    }
  }
}

void DgGroup::TimeStepBits(unsigned long long *dt_min_bits, hipStream_t stream) {
  Check(LaunchDgTimeStep(cell_type, degree, ct0.ptr, geo.ptr, tables_d.ptr, n_cells, dt_min_bits, stream),
        "DG time-step launch");
}

void DgGroup::Approximate(const double *values, hipStream_t stream) {
  DeviceArray<double> v;
  v.Allocate(size_t(n_cells) * nq * 5);
  Check(hipMemcpyAsync(v.ptr, values, size_t(n_cells) * nq * 5 * sizeof(double), hipMemcpyHostToDevice, stream),
        "approximate upload");
  Check(LaunchDgApproximate(degree, v.ptr, basis.ptr, geo.ptr, tables_d.ptr, nq, n_cells, u0.ptr, stream),
        "DG approximate launch");
  SolutionChanged(stream);
  Check(hipStreamSynchronize(stream), "approximate sync");  // before the values buffer is freed
}

void DgGroup::Faces(int64_t *holder, int64_t *sharer, int32_t *holder_face, int32_t *sharer_face) const {
  for (size_t f = 0; f < faces.size(); ++f) {
    if (holder) holder[f] = faces[f].holder;
    if (sharer) sharer[f] = faces[f].sharer;
    if (holder_face) holder_face[f] = faces[f].holder_face;
    if (sharer_face) sharer_face[f] = faces[f].sharer_face;
  }
}

void DgGroup::Evaluate(const double *column, int n_local, const double *local, double *xyz,
                       double *values) const {
  const bool hex = cell_type == MCFD_HEXA_8;
  const int packed = n * (n + 1) / 2;
  for (int64_t c = 0; c < n_cells; ++c) {
    const double *g = h_geo.data() + c * kDgGeoStride;  // J rows, |det J|, length, centre
    const double *s = h_basis.data() + c * packed;
    const double *uc = column + c * n * 5;
    for (int p = 0; p < n_local; ++p) {
      double eta[3];
      for (int d = 0; d < 3; ++d) eta[d] = local[3 * p + d] - (hex ? 0.0 : 0.25);
      double y[3];
      for (int d = 0; d < 3; ++d) y[d] = eta[0] * g[d] + eta[1] * g[3 + d] + eta[2] * g[6 + d];
      if (xyz)
        for (int d = 0; d < 3; ++d) xyz[(c * n_local + p) * 3 + d] = g[11 + d] + y[d];
      if (!values) continue;
      // Taylor monomials (basis/taylor.hpp), then the orthonormal basis phi = S T
      const double t[10] = { 1.0, y[0], y[1], y[2], y[0] * y[0], y[0] * y[1], y[0] * y[2],
                             y[1] * y[1], y[1] * y[2], y[2] * y[2] };
      double *v = values + (c * n_local + p) * 5;
      for (int k = 0; k < 5; ++k) v[k] = 0;
      for (int i = 0, e = 0; i < n; ++i) {
        double phi = 0;
        for (int jj = 0; jj <= i; ++jj, ++e) phi += s[e] * t[jj];
        for (int k = 0; k < 5; ++k) v[k] += uc[i * 5 + k] * phi;
      }
    }
  }
}

}  // namespace mcfd
