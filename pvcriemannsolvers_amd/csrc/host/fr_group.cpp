// FrHexGroup: host construction and device upload of one FR mesh.
#include "fr_group.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>

#include "tables.hpp"

namespace mcfd {

namespace {

// Inverse of an orientation code (see ApplyOrient in hex_connectivity.hpp).
int InverseOrient(int code) {
  const int q = 4;
  for (int c = 0; c < 8; ++c) {
    bool ok = true;
    for (int a = 0; a < q && ok; ++a)
      for (int b = 0; b < q && ok; ++b) {
        int a1, b1, a2, b2;
        ApplyOrient(code, q, a, b, &a1, &b1);
        ApplyOrient(c, q, a1, b1, &a2, &b2);
        ok = (a2 == a && b2 == b);
      }
    if (ok) return c;
  }
  return -1;
}

// J* = det(J) J^{-1}, the adjugate of J (rows d/dxi, d/deta, d/dzeta), row-major
// so that F~_d = sum_e F_e g[3 e + d]; returns det(J).
double Adjugate(const double *j, double *g) {
  g[0] = j[4] * j[8] - j[5] * j[7]; g[1] = j[2] * j[7] - j[1] * j[8]; g[2] = j[1] * j[5] - j[2] * j[4];
  g[3] = j[5] * j[6] - j[3] * j[8]; g[4] = j[0] * j[8] - j[2] * j[6]; g[5] = j[2] * j[3] - j[0] * j[5];
  g[6] = j[3] * j[7] - j[4] * j[6]; g[7] = j[1] * j[6] - j[0] * j[7]; g[8] = j[0] * j[4] - j[1] * j[3];
  return j[0] * g[0] + j[1] * g[3] + j[2] * g[6];
}

// Local coordinates of flux point (a, b) of local face f (the end of the
// solution-point line through it, fr/general.hpp:103-205).
void FacePointLocal(int f, const double *x1d, int a, int b, double *s) {
  const int dir = kHexFaceNormalDir[f];
  s[dir] = kHexFaceSide[f] ? 1.0 : -1.0;
  if (dir == 2) { s[0] = x1d[a]; s[1] = x1d[b]; }
  else if (dir == 1) { s[0] = x1d[a]; s[2] = x1d[b]; }
  else { s[1] = x1d[a]; s[2] = x1d[b]; }
}

}  // namespace

// Geometry of a mesh with trilinear cells (polynomial/hexahedron.hpp:172-239
// keeps J* and det(J) per node; fr/general.hpp:103-205 the flux point's
// J* column; the face integrator the frame of each face point):
//   geo_node [cell][node]  J* (9), 1 / det(J) at the node;
//   fscale [cell][face][a * Q + b]  +-|J*.col(normal dir)| at the flux point;
//   frames [face][a * Q + b]  n, t of the face map at the point, indexed by the
//     holder's in-face (a, b);
//   geo [cell]  only the length cbrt(volume) / 2, volume = sum of w det(J).
void FrHexGroup::BuildCurved(const mcfd_scheme_desc &) {
  const int64_t nc = n_cells;
  const int nfp = q * q;
  diag_metric = false;
  h_geo.assign(nc * kGeoStride, 0.0);
  h_geo_id.resize(nc);
  h_geo_node.assign(nc * np * kNodeGeoStride, 0.0);
  h_fscale.assign(nc * 6 * nfp, 0.0);
  corner0.resize(nc);
  for (int64_t c = 0; c < nc; ++c) {
    const double *xv = conn.vertices[c].data();
    double volume = 0;
    for (int i = 0; i < q; ++i)
      for (int jj = 0; jj < q; ++jj)
        for (int k = 0; k < q; ++k) {
          const double s[3] = { x1d[i], x1d[jj], x1d[k] };
          double jac[9];
          TrilinearJacobian(xv, s, jac);
          double *g = h_geo_node.data() + (c * np + (i * q + jj) * q + k) * kNodeGeoStride;
          const double det = Adjugate(jac, g);
          if (!(det > 0)) throw std::invalid_argument("inverted trilinear hexahedron");
          g[9] = 1.0 / det;
          volume += w1d[i] * w1d[jj] * w1d[k] * det;
        }
    h_geo[c * kGeoStride + 16] = std::cbrt(volume) * 0.5;  // mesh/part.hpp:230
    h_geo_id[c] = (int32_t)c;
    for (int f = 0; f < 6; ++f) {
      const int dir = kHexFaceNormalDir[f];
      for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b) {
          double s[3], jac[9], g[9];
          FacePointLocal(f, x1d.data(), a, b, s);
          TrilinearJacobian(xv, s, jac);
          Adjugate(jac, g);
          const double n = std::sqrt(g[dir] * g[dir] + g[3 + dir] * g[3 + dir] + g[6 + dir] * g[6 + dir]);
          h_fscale[(c * 6 + f) * nfp + a * q + b] = kHexFaceSide[f] ? n : -n;
        }
    }
    corner0[c] = { xv[0], xv[1], xv[2] };
  }
  if ((int64_t)conn.faces.size() >= (int64_t(1) << 24))
    throw Failure{ MCFD_ENOTSUP, "too many faces for the per-point frames of a trilinear mesh" };
  h_frames.assign(conn.faces.size() * nfp * kFrameStride, 0.0);
  h_nbr.assign(6 * nc, INT32_MIN);
  h_finfo.assign(6 * nc, 0);
  for (size_t fi = 0; fi < conn.faces.size(); ++fi) {
    const auto &face = conn.faces[fi];
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) {
        int a, b;
        ApplyOrient(face.quad_to_holder, q, i, j, &a, &b);
        FaceFrameAt(face.corners, x1d[i], x1d[j], h_frames.data() + (fi * nfp + a * q + b) * kFrameStride);
      }
    const int32_t fid = (int32_t)fi;
    const int64_t hs = face.holder * 6 + face.holder_face;
    if (face.sharer >= 0) {
      const int64_t ss = face.sharer * 6 + face.sharer_face;
      h_nbr[hs] = (int32_t)face.sharer;
      h_finfo[hs] = (fid << 7) | (1 << 6) | (face.orient << 3) | face.sharer_face;
      h_nbr[ss] = (int32_t)face.holder;
      h_finfo[ss] = (fid << 7) | (InverseOrient(face.orient) << 3) | face.holder_face;
    } else {
      h_nbr[hs] = -(1 + face.tag);
      h_finfo[hs] = (fid << 7) | (1 << 6);
    }
  }
  for (auto v : h_nbr)
    if (v == INT32_MIN) throw std::invalid_argument("cell face without a face record");
  h_bface_of.assign(6 * nc, -1);
  for (size_t f = conn.n_interior; f < conn.faces.size(); ++f)
    h_bface_of[conn.faces[f].holder * 6 + conn.faces[f].holder_face] = (int32_t)(f - conn.n_interior);
}

void FrHexGroup::Build(const mcfd_mesh_desc &mesh, const mcfd_scheme_desc &scheme,
                       const std::vector<int32_t> &bc_kinds, const std::vector<double> &bc_states) {
  if (scheme.degree < 1 || scheme.degree > 5) throw std::invalid_argument("degree must be 1..5");
  if (scheme.riemann < 0 || scheme.riemann > 2) throw std::invalid_argument("unknown Riemann solver");
  if (scheme.precision != MCFD_PRECISION_FP64 && scheme.precision != MCFD_PRECISION_MIXED)
    throw std::invalid_argument("unknown precision form");
  degree = scheme.degree;
  riemann = scheme.riemann;
  mixed = scheme.precision == MCFD_PRECISION_MIXED;
  q = degree + 1;
  np = q * q * q;
  n_cells = mesh.n_cells;
  const int n_bcs = (int)bc_kinds.size();
  for (int64_t b = 0; b < mesh.n_bfaces; ++b)
    if (mesh.bface_tags[b] >= n_bcs) throw std::invalid_argument("boundary tag without a BC entry");
  for (auto k : bc_kinds)
    if (k < 0 || k > 5) throw std::invalid_argument("unknown BC kind");
  h_bc_kind = bc_kinds;
  h_bc_state = bc_states;
  if (h_bc_kind.empty()) {  // keep the device arrays non-empty
    h_bc_kind.push_back(1);
    h_bc_state.assign(5, 0.0);
  }
  conn = BuildHexConnectivity(mesh);

  // line tables
  x1d.resize(q);
  w1d.resize(q);
  GaussRule(scheme.nodes == MCFD_NODES_LOBATTO, q, x1d.data(), w1d.data());
  std::memset(&lt, 0, sizeof(lt));
  for (int i = 0; i < q; ++i)
    for (int m = 0; m < q; ++m) lt.d[i][m] = LagrangeDerivative(q, x1d.data(), m, x1d[i]);
  for (int m = 0; m < q; ++m) {
    lt.lend[0][m] = LagrangeValue(q, x1d.data(), m, -1.0);
    lt.lend[1][m] = LagrangeValue(q, x1d.data(), m, +1.0);
    lt.gp[0][m] = CorrectionLeftDerivative(degree, scheme.c_next, x1d[m]);
    lt.gp[1][m] = CorrectionRightDerivative(degree, scheme.c_next, x1d[m]);
  }

  const int64_t nc = mesh.n_cells;
  curved = false;
  for (auto a : conn.affine)
    if (!a) curved = true;
  if (curved) {
    if (mixed) throw Failure{ MCFD_ENOTSUP, "the mixed-precision form supports affine hexahedra only" };
    BuildCurved(scheme);
    return;
  }

  // geometry records, deduplicated bitwise
  std::map<std::array<double, kGeoStride>, int32_t> geo_ids;
  h_geo.clear();
  h_geo_id.assign(nc, 0);
  corner0.resize(nc);
  for (int64_t c = 0; c < nc; ++c) {
    const auto &j = conn.jac[c];
    double d = conn.det[c];
    // J* = det(J) J^{-1} is the adjugate of J
    std::array<double, kGeoStride> g{};
    g[0] = j[4] * j[8] - j[5] * j[7]; g[1] = j[2] * j[7] - j[1] * j[8]; g[2] = j[1] * j[5] - j[2] * j[4];
    g[3] = j[5] * j[6] - j[3] * j[8]; g[4] = j[0] * j[8] - j[2] * j[6]; g[5] = j[2] * j[3] - j[0] * j[5];
    g[6] = j[3] * j[7] - j[4] * j[6]; g[7] = j[1] * j[6] - j[0] * j[7]; g[8] = j[0] * j[4] - j[1] * j[3];
    g[9] = 1.0 / d;
    for (int f = 0; f < 6; ++f) {
      int dir = kHexFaceNormalDir[f];
      double n = std::sqrt(g[dir] * g[dir] + g[3 + dir] * g[3 + dir] + g[6 + dir] * g[6 + dir]);
      g[10 + f] = kHexFaceSide[f] ? n : -n;
    }
    g[16] = std::cbrt(8 * d) * 0.5;  // mesh/part.hpp:230, volume = 8 det(J)
    auto it = geo_ids.find(g);
    if (it == geo_ids.end()) {
      it = geo_ids.emplace(g, (int32_t)geo_ids.size()).first;
      h_geo.insert(h_geo.end(), g.begin(), g.end());
    }
    h_geo_id[c] = it->second;
    int64_t n0 = mesh.cell_nodes[8 * c];
    corner0[c] = { mesh.coords[3 * n0], mesh.coords[3 * n0 + 1], mesh.coords[3 * n0 + 2] };
  }
  diag_metric = true;
  for (size_t r = 0; r < h_geo.size(); r += kGeoStride)
    for (int e = 0; e < 9; ++e)
      if (e % 4 != 0 && h_geo[r + e] != 0.0) diag_metric = false;

  // per-cell face records
  std::map<std::array<double, kFrameStride>, int32_t> frame_ids;
  h_frames.clear();
  h_nbr.assign(6 * nc, INT32_MIN);
  h_finfo.assign(6 * nc, 0);
  for (const auto &face : conn.faces) {
    std::array<double, kFrameStride> fr;
    std::copy(face.frame, face.frame + 6, fr.begin());
    auto it = frame_ids.find(fr);
    if (it == frame_ids.end()) {
      it = frame_ids.emplace(fr, (int32_t)frame_ids.size()).first;
      h_frames.insert(h_frames.end(), fr.begin(), fr.end());
    }
    int32_t fid = it->second;
    if (fid >= (1 << 24)) throw Failure{ MCFD_ENOTSUP, "too many distinct face frames" };
    int64_t hs = face.holder * 6 + face.holder_face;
    if (face.sharer >= 0) {
      int64_t ss = face.sharer * 6 + face.sharer_face;
      h_nbr[hs] = (int32_t)face.sharer;
      h_finfo[hs] = (fid << 7) | (1 << 6) | (face.orient << 3) | face.sharer_face;
      h_nbr[ss] = (int32_t)face.holder;
      h_finfo[ss] = (fid << 7) | (InverseOrient(face.orient) << 3) | face.holder_face;
    } else {
      h_nbr[hs] = -(1 + face.tag);
      h_finfo[hs] = (fid << 7) | (1 << 6);
    }
  }
  for (auto v : h_nbr)
    if (v == INT32_MIN) throw std::invalid_argument("cell face without a face record");
  h_bface_of.assign(6 * nc, -1);
  for (size_t f = conn.n_interior; f < conn.faces.size(); ++f)
    h_bface_of[conn.faces[f].holder * 6 + conn.faces[f].holder_face] = (int32_t)(f - conn.n_interior);
}

void FrHexGroup::BoundaryPoints(std::vector<int32_t> *tags, std::vector<double> *xyz) const {
  const int64_t nb = (int64_t)conn.faces.size() - conn.n_interior;
  tags->assign(nb, 0);
  xyz->assign(nb * q * q * 3, 0.0);
  for (int64_t b = 0; b < nb; ++b) {
    const auto &face = conn.faces[conn.n_interior + b];
    (*tags)[b] = face.tag;
    const int lf = face.holder_face, dir = kHexFaceNormalDir[lf];
    const double end = kHexFaceSide[lf] ? 1.0 : -1.0;
    const auto &j = conn.jac[face.holder];
    const auto &x0 = corner0[face.holder];
    for (int ia = 0; ia < q; ++ia)
      for (int ib = 0; ib < q; ++ib) {
        // reference point of flux point (a, b): the line end of fr/general.hpp:103-205
        double s[3];
        s[dir] = end;
        if (dir == 2) { s[0] = x1d[ia]; s[1] = x1d[ib]; }
        else if (dir == 1) { s[0] = x1d[ia]; s[2] = x1d[ib]; }
        else { s[1] = x1d[ia]; s[2] = x1d[ib]; }
        double *p = xyz->data() + ((b * q + ia) * q + ib) * 3;
        if (curved) {
          TrilinearPoint(conn.vertices[face.holder].data(), s, p);
          continue;
        }
        for (int d = 0; d < 3; ++d)
          p[d] = x0[d] + (s[0] + 1) * j[d] + (s[1] + 1) * j[3 + d] + (s[2] + 1) * j[6 + d];
      }
  }
}

void FrHexGroup::Upload(hipStream_t stream) {
  size_t cs = ColumnSize();
  u0.Allocate(cs); u1.Allocate(cs); u2.Allocate(cs); res.Allocate(cs);
  Check(hipMemsetAsync(u0.ptr, 0, cs * sizeof(double), stream), "memset");
  nbr.Upload(h_nbr, stream);
  finfo.Upload(h_finfo, stream);
  geo_id.Upload(h_geo_id, stream);
  geo.Upload(h_geo, stream);
  frames.Upload(h_frames, stream);
  if (curved) {
    geo_node.Upload(h_geo_node, stream);
    fscale.Upload(h_fscale, stream);
  }
  bc_kind.Upload(h_bc_kind, stream);
  bc_state.Upload(h_bc_state, stream);
  if (!h_cell_case.empty()) cell_case.Upload(h_cell_case, stream);
  if (!h_tile_order.empty()) tile_order.Upload(h_tile_order, stream);
  bface_of.Upload(h_bface_of, stream);
  if (riemann == MCFD_RIEMANN_EXACT) {
    const size_t n_star = size_t(n_cells) * 6 * q * q;
    exact_star.Allocate(n_star);
    Check(hipMemsetAsync(exact_star.ptr, 0, n_star, stream), "memset");
  }
}

void FrHexGroup::BuildTileOrder(int tile_cells) {
  h_tile_order.clear();
  if (tile_cells < 1 || n_cells == 0) return;
  const int64_t n_tiles = (n_cells + tile_cells - 1) / tile_cells;
  std::vector<std::array<double, 3>> centre(n_tiles, { 0, 0, 0 });
  double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
  for (int64_t t = 0; t < n_tiles; ++t) {
    const int64_t c0 = t * tile_cells, c1 = std::min<int64_t>(n_cells, c0 + tile_cells);
    for (int64_t c = c0; c < c1; ++c)
      for (int d = 0; d < 3; ++d) centre[t][d] += corner0[c][d] / double(c1 - c0);
    for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], centre[t][d]); hi[d] = std::max(hi[d], centre[t][d]); }
  }
  auto spread = [](uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  std::vector<std::pair<uint64_t, int32_t>> key(n_tiles);
  for (int64_t t = 0; t < n_tiles; ++t) {
    uint64_t code = 0;
    for (int d = 0; d < 3; ++d) {
      const double span = hi[d] - lo[d];
      const uint64_t q = span > 0 ? (uint64_t)((centre[t][d] - lo[d]) / span * 2097151.0) : 0;
      code |= spread(q) << d;
    }
    key[t] = { code, (int32_t)t };
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const auto &a, const auto &b) { return a.first < b.first; });
  h_tile_order.resize(n_tiles);
  for (int64_t t = 0; t < n_tiles; ++t) h_tile_order[t] = key[t].second;
}

FrHexArgs FrHexGroup::Args(int stage, double dt) const {
  FrHexArgs a;
  a.u0 = u0.ptr;
  a.nbr = nbr.ptr; a.finfo = finfo.ptr; a.geo_id = geo_id.ptr;
  a.geo = geo.ptr; a.frames = frames.ptr;
  a.bc_kind = bc_kind.ptr; a.bc_state = bc_state.ptr;
  a.cell_case = cell_case.ptr;
  a.case_dt = cell_case.ptr ? case_dt.ptr : nullptr;
  a.n_cells = n_cells;
  a.dt = dt;
  a.stage = stage;
  a.xcd_remap = 1;
  a.tile_order = tile_order.ptr;
  a.dt_min_bits = nullptr;
  a.exact_star = exact_star.ptr;
  a.bc_point_state = nullptr;  // set by the caller when boundary functions are in use
  a.bface_of = bface_of.ptr;
  a.diag_metric = diag_metric ? 1 : 0;
  a.mixed = mixed ? 1 : 0;
  a.curved = curved ? 1 : 0;
  a.geo_node = geo_node.ptr;
  a.fscale = fscale.ptr;
  switch (stage) {
  case 0: a.us = u0.ptr; a.out = res.ptr; break;
  case 1: a.us = u0.ptr; a.out = u1.ptr; break;
  case 2: a.us = u1.ptr; a.out = u2.ptr; break;
  default: a.us = u2.ptr; a.out = u0.ptr; break;
  }
  return a;
}

void FrHexGroup::SamplePoints(double *xyz) const {
  for (int64_t c = 0; c < n_cells; ++c) {
    const auto &j = conn.jac[c];
    const auto &x0 = corner0[c];
    for (int i = 0; i < q; ++i)
      for (int jj = 0; jj < q; ++jj)
        for (int k = 0; k < q; ++k) {
          double *p = xyz + ((c * np) + (i * q + jj) * q + k) * 3;
          if (curved) {
            const double l[3] = { x1d[i], x1d[jj], x1d[k] };
            TrilinearPoint(conn.vertices[c].data(), l, p);
            continue;
          }
          double s[3] = { x1d[i] + 1, x1d[jj] + 1, x1d[k] + 1 };
          for (int d = 0; d < 3; ++d) p[d] = x0[d] + s[0] * j[d] + s[1] * j[3 + d] + s[2] * j[6 + d];
        }
  }
}

void FrHexGroup::Evaluate(const double *column, int n_local, const double *local, double *xyz,
                          double *values) const {
  std::vector<double> la(q), lb(q), lc(q);
  for (int64_t c = 0; c < n_cells; ++c) {
    const double *uc = column + c * np * 5;
    for (int p = 0; p < n_local; ++p) {
      const double *l = local + 3 * p;
      if (xyz) {
        double *x = xyz + (c * n_local + p) * 3;
        if (curved) {
          TrilinearPoint(conn.vertices[c].data(), l, x);
        } else {
          const auto &j = conn.jac[c];
          const auto &x0 = corner0[c];
          for (int d = 0; d < 3; ++d)
            x[d] = x0[d] + (l[0] + 1) * j[d] + (l[1] + 1) * j[3 + d] + (l[2] + 1) * j[6 + d];
        }
      }
      if (!values) continue;
      for (int m = 0; m < q; ++m) {
        la[m] = LagrangeValue(q, x1d.data(), m, l[0]);
        lb[m] = LagrangeValue(q, x1d.data(), m, l[1]);
        lc[m] = LagrangeValue(q, x1d.data(), m, l[2]);
      }
      double *v = values + (c * n_local + p) * 5;
      for (int k = 0; k < 5; ++k) v[k] = 0;
      for (int i = 0; i < q; ++i)
        for (int jj = 0; jj < q; ++jj)
          for (int kk = 0; kk < q; ++kk) {
            const double w = la[i] * lb[jj] * lc[kk];
            const double *u = uc + ((i * q + jj) * q + kk) * 5;
            for (int k = 0; k < 5; ++k) v[k] += w * u[k];
          }
    }
  }
}

std::vector<double> FrHexGroup::NodeDeterminants() const {
  std::vector<double> det(size_t(n_cells) * np);
  for (int64_t c = 0; c < n_cells; ++c)
    for (int i = 0; i < q; ++i)
      for (int jj = 0; jj < q; ++jj)
        for (int k = 0; k < q; ++k) {
          double &out = det[c * np + (i * q + jj) * q + k];
          if (!curved) {
            out = conn.det[c];
            continue;
          }
          const double s[3] = { x1d[i], x1d[jj], x1d[k] };
          double jac[9], g[9];
          TrilinearJacobian(conn.vertices[c].data(), s, jac);
          out = Adjugate(jac, g);
        }
  return det;
}

void FrHexGroup::Faces(int64_t *holder, int64_t *sharer, int32_t *holder_face, int32_t *sharer_face,
                       int32_t *fp_ijk) const {
  const int nfp = q * q;
  for (size_t f = 0; f < conn.faces.size(); ++f) {
    const auto &face = conn.faces[f];
    if (holder) holder[f] = face.holder;
    if (sharer) sharer[f] = face.sharer;
    if (holder_face) holder_face[f] = face.holder_face;
    if (sharer_face) sharer_face[f] = face.sharer_face;
    if (!fp_ijk) continue;
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) {
        int a, b;
        ApplyOrient(face.quad_to_holder, q, i, j, &a, &b);
        int fh = face.holder_face;
        int32_t *o = fp_ijk + (f * nfp + i * q + j) * 2;
        o[0] = HexLineNode(q, fh, a, b, kHexFaceSide[fh] ? q - 1 : 0);
        o[1] = -1;
        if (face.sharer >= 0) {
          int a2, b2, fs = face.sharer_face;
          ApplyOrient(face.orient, q, a, b, &a2, &b2);
          o[1] = HexLineNode(q, fs, a2, b2, kHexFaceSide[fs] ? q - 1 : 0);
        }
      }
  }
}

}  // namespace mcfd
