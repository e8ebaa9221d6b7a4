// Connectivity builder for conforming hexahedral meshes (see the header for
// the reference semantics it reproduces).
#include "hex_connectivity.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>

namespace mcfd {

namespace {

// CGNS HEXA_8 corner coordinates in (xi, eta, zeta) and the corner lists of
// its six faces (CGNS element conventions; reference tables at
// coordinate/hexahedron.hpp:353-359,379-389).
constexpr int kCorner[8][3] = {
  { -1, -1, -1 }, { +1, -1, -1 }, { +1, +1, -1 }, { -1, +1, -1 },
  { -1, -1, +1 }, { +1, -1, +1 }, { +1, +1, +1 }, { -1, +1, +1 },
};
constexpr int kFaceCorners[6][4] = {
  { 0, 3, 2, 1 }, { 0, 1, 5, 4 }, { 1, 2, 6, 5 },
  { 2, 3, 7, 6 }, { 0, 4, 7, 3 }, { 4, 5, 6, 7 },
};
// CGNS QUAD_4 corner coordinates (coordinate/quadrangle.hpp:136-139).
constexpr int kQuadCorner[4][2] = { { -1, -1 }, { +1, -1 }, { +1, +1 }, { -1, +1 } };

using Vec = std::array<double, 3>;

Vec Sub(Vec a, const Vec &b) { for (int i = 0; i < 3; ++i) a[i] -= b[i]; return a; }
Vec Add(Vec a, const Vec &b) { for (int i = 0; i < 3; ++i) a[i] += b[i]; return a; }
double Norm(const Vec &a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
Vec Cross(const Vec &a, const Vec &b) {
  return { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
}
Vec Unit(const Vec &a) { double n = Norm(a); return { a[0] / n, a[1] / n, a[2] / n }; }

void InFace(int f, const int *loc, int *s, int *t) {
  switch (f) {
  case 0: case 5: *s = loc[0]; *t = loc[1]; break;
  case 1: case 3: *s = loc[0]; *t = loc[2]; break;
  default: *s = loc[1]; *t = loc[2]; break;
  }
}

// Corner index of cell-local face f at in-face coordinates (s, t).
int CornerAt(int f, int s, int t) {
  int loc[3];
  int dir = kHexFaceNormalDir[f];
  loc[dir] = kHexFaceSide[f] ? 1 : -1;
  int o0 = dir == 0 ? 1 : 0, o1 = dir == 2 ? 1 : 2;
  loc[o0] = s;
  loc[o1] = t;
  for (int c = 0; c < 8; ++c)
    if (kCorner[c][0] == loc[0] && kCorner[c][1] == loc[1] && kCorner[c][2] == loc[2]) return c;
  return -1;
}

// Orientation code of the affine map (s, t) -> (s', t') given the images of
// (-1, -1) and (+1, -1).
int OrientFrom(int s0, int t0, int s1, int t1) {
  bool swap = (s1 == s0);
  return (swap ? 1 : 0) | (s0 > 0 ? 2 : 0) | (t0 > 0 ? 4 : 0);
}

struct FaceKey {
  std::array<int64_t, 4> n;
  int64_t slot;  // cell * 6 + local face
  bool operator<(const FaceKey &o) const { return n < o.n || (n == o.n && slot < o.slot); }
};

}  // namespace

// coordinate/hexahedron.hpp:397-526 (Hexahedron8): x = sum_i N_i(xi) x_i with
// N_i = (1 + a_i xi)(1 + b_i eta)(1 + c_i zeta) / 8 over the CGNS corners.
void TrilinearPoint(const double *xv, const double *xi, double *x) {
  x[0] = x[1] = x[2] = 0;
  for (int i = 0; i < 8; ++i) {
    const double v = (1 + kCorner[i][0] * xi[0]) * (1 + kCorner[i][1] * xi[1]) * (1 + kCorner[i][2] * xi[2]) / 8;
    for (int d = 0; d < 3; ++d) x[d] += xv[3 * i + d] * v;
  }
}

void TrilinearJacobian(const double *xv, const double *xi, double *j) {
  for (int e = 0; e < 9; ++e) j[e] = 0;
  for (int i = 0; i < 8; ++i) {
    const double fx = 1 + kCorner[i][0] * xi[0], fy = 1 + kCorner[i][1] * xi[1], fz = 1 + kCorner[i][2] * xi[2];
    const double g[3] = { kCorner[i][0] * fy * fz / 8, kCorner[i][1] * fz * fx / 8, kCorner[i][2] * fx * fy / 8 };
    for (int r = 0; r < 3; ++r)
      for (int d = 0; d < 3; ++d) j[3 * r + d] += g[r] * xv[3 * i + d];
  }
}

void FaceFrameAt(const double *xq, double a, double b, double *frame) {
  Vec ra{ 0, 0, 0 }, rb{ 0, 0, 0 };
  for (int i = 0; i < 4; ++i)
    for (int d = 0; d < 3; ++d) {
      ra[d] += kQuadCorner[i][0] * (1 + kQuadCorner[i][1] * b) / 4 * xq[3 * i + d];
      rb[d] += kQuadCorner[i][1] * (1 + kQuadCorner[i][0] * a) / 4 * xq[3 * i + d];
    }
  Vec n = Unit(Cross(ra, rb)), t = Unit(ra);
  for (int d = 0; d < 3; ++d) { frame[d] = n[d]; frame[3 + d] = t[d]; }
}

HexConnectivity BuildHexConnectivity(const mcfd_mesh_desc &mesh) {
  if (mesh.cell_type != MCFD_HEXA_8) throw std::invalid_argument("hexahedral mesh expected");
  if (mesh.n_cells <= 0 || mesh.n_cells > (int64_t(1) << 30))
    throw std::invalid_argument("cell count out of range");
  const int64_t nc = mesh.n_cells;
  auto node = [&](int64_t id) -> Vec {
    if (id < 0 || id >= mesh.n_nodes) throw std::invalid_argument("node id out of range");
    return { mesh.coords[3 * id], mesh.coords[3 * id + 1], mesh.coords[3 * id + 2] };
  };
  auto cell_node = [&](int64_t c, int i) { return mesh.cell_nodes[8 * c + i]; };

  HexConnectivity out;
  out.n_cells = nc;
  out.jac.resize(nc);
  out.det.resize(nc);
  out.affine.resize(nc);
  out.vertices.resize(nc);
  for (int64_t c = 0; c < nc; ++c) {
    Vec x[8];
    for (int i = 0; i < 8; ++i) x[i] = node(cell_node(c, i));
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d) out.vertices[c][3 * i + d] = x[i][d];
    Vec e1 = Sub(x[1], x[0]), e2 = Sub(x[3], x[0]), e3 = Sub(x[4], x[0]);
    double h = std::max({ Norm(e1), Norm(e2), Norm(e3) });
    const Vec expect[8] = { x[0], x[1], Add(x[1], e2), x[3], x[4], Add(x[4], e1),
                            Add(Add(x[4], e1), e2), Add(x[4], e2) };
    bool affine = true;
    for (int i = 0; i < 8; ++i)
      if (Norm(Sub(expect[i], x[i])) > 1e-10 * h) affine = false;
    out.affine[c] = affine ? 1 : 0;
    auto &j = out.jac[c];
    if (affine) {
      for (int d = 0; d < 3; ++d) { j[d] = e1[d] / 2; j[3 + d] = e2[d] / 2; j[6 + d] = e3[d] / 2; }
      out.det[c] = j[0] * (j[4] * j[8] - j[5] * j[7]) - j[1] * (j[3] * j[8] - j[5] * j[6])
                 + j[2] * (j[3] * j[7] - j[4] * j[6]);
      if (!(out.det[c] > 0)) throw std::invalid_argument("inverted or degenerate hexahedron");
    } else {
      // trilinear cell: the Jacobian varies; its centre value is kept here and
      // the node and face-point values are evaluated by TrilinearJacobian
      const double zero[3] = { 0, 0, 0 };
      TrilinearJacobian(out.vertices[c].data(), zero, j.data());
      out.det[c] = j[0] * (j[4] * j[8] - j[5] * j[7]) - j[1] * (j[3] * j[8] - j[5] * j[6])
                 + j[2] * (j[3] * j[7] - j[4] * j[6]);
      if (!(out.det[c] > 0)) throw std::invalid_argument("inverted or degenerate hexahedron");
    }
  }

  // Face matching by sorted corner ids.
  std::vector<FaceKey> keys(6 * nc);
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < 6; ++f) {
      auto &k = keys[6 * c + f];
      for (int i = 0; i < 4; ++i) k.n[i] = cell_node(c, kFaceCorners[f][i]);
      std::sort(k.n.begin(), k.n.end());
      k.slot = 6 * c + f;
    }
  std::sort(keys.begin(), keys.end());
  struct Pair { int64_t a, b; Vec shift; };  // slots, sharer = holder + shift
  std::vector<Pair> pairs;
  std::vector<FaceKey> open;  // faces with one cell
  for (size_t i = 0; i < keys.size();) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].n == keys[i].n) ++j;
    if (j - i == 2) pairs.push_back({ keys[i].slot, keys[i + 1].slot, { 0, 0, 0 } });
    else if (j - i == 1) open.push_back(keys[i]);
    else throw std::invalid_argument("non-manifold face shared by more than two cells");
    i = j;
  }

  // Boundary faces: find their cell face among the open ones.
  struct Bnd { int64_t slot; int32_t tag; };
  std::vector<Bnd> bnd(mesh.n_bfaces);
  std::vector<char> used(open.size(), 0);
  for (int64_t b = 0; b < mesh.n_bfaces; ++b) {
    FaceKey k;
    for (int i = 0; i < 4; ++i) k.n[i] = mesh.bface_nodes[4 * b + i];
    std::sort(k.n.begin(), k.n.end());
    k.slot = -1;
    auto it = std::lower_bound(open.begin(), open.end(), k);
    if (it == open.end() || it->n != k.n) throw std::invalid_argument("boundary face is not a cell face");
    if (used[it - open.begin()]++) throw std::invalid_argument("duplicate boundary face");
    bnd[b] = { it->slot, mesh.bface_tags[b] };
  }
  for (size_t i = 0; i < open.size(); ++i)
    if (!used[i]) throw std::invalid_argument("cell face neither shared nor on the boundary");

  auto face_center = [&](int64_t slot) {
    Vec s{ 0, 0, 0 };
    for (int i = 0; i < 4; ++i) s = Add(s, node(cell_node(slot / 6, kFaceCorners[slot % 6][i])));
    return Vec{ s[0] / 4, s[1] / 4, s[2] / 4 };
  };
  // Periodic pairing by translated centres.
  {
    double span = 0;
    for (int d = 0; d < 3; ++d) span = std::max(span, mesh.period[d]);
    for (int64_t i = 0; i < mesh.n_nodes; ++i)
      for (int d = 0; d < 3; ++d) span = std::max(span, std::abs(mesh.coords[3 * i + d]));
    const double h = 1e-7 * (span > 0 ? span : 1);
    using QKey = std::tuple<int64_t, int64_t, int64_t>;
    std::map<QKey, size_t> by_center;
    std::vector<size_t> periodic;
    std::vector<Vec> centers;
    for (size_t b = 0; b < bnd.size(); ++b) {
      if (bnd[b].tag >= 0) continue;
      Vec c = face_center(bnd[b].slot);
      by_center[{ std::llround(c[0] / h), std::llround(c[1] / h), std::llround(c[2] / h) }] = periodic.size();
      periodic.push_back(b);
      centers.push_back(c);
    }
    std::vector<int64_t> partner(periodic.size(), -1);
    std::vector<Vec> partner_shift(periodic.size());
    for (size_t p = 0; p < periodic.size(); ++p) {
      if (partner[p] >= 0) continue;
      bool found = false;
      for (int d = 0; d < 3 && !found; ++d) {
        if (!(mesh.period[d] > 0)) continue;
        for (int sgn = -1; sgn <= 1 && !found; sgn += 2) {
          Vec s{ 0, 0, 0 };
          s[d] = sgn * mesh.period[d];
          Vec t = Add(centers[p], s);
          double g[3] = { t[0] / h, t[1] / h, t[2] / h };
          for (int m = 0; m < 8 && !found; ++m) {
            QKey q{ (m & 1) ? (int64_t)std::ceil(g[0]) : (int64_t)std::floor(g[0]),
                    (m & 2) ? (int64_t)std::ceil(g[1]) : (int64_t)std::floor(g[1]),
                    (m & 4) ? (int64_t)std::ceil(g[2]) : (int64_t)std::floor(g[2]) };
            auto it = by_center.find(q);
            if (it == by_center.end() || it->second == p) continue;
            size_t o = it->second;
            if (partner[o] >= 0) continue;
            if (Norm(Sub(centers[o], t)) > 10 * h) continue;
            partner[p] = o; partner[o] = p;
            partner_shift[p] = s;
            found = true;
          }
        }
      }
      if (!found) throw std::invalid_argument("periodic face without a partner across the box");
      int64_t sa = bnd[periodic[p]].slot, sb = bnd[periodic[partner[p]]].slot;
      pairs.push_back({ sa, sb, partner_shift[p] });
    }
  }

  // Interior faces: holder = lower cell id; sorted by (holder, sharer, face).
  for (auto &p : pairs) {
    if (p.b / 6 < p.a / 6) {
      std::swap(p.a, p.b);
      for (auto &x : p.shift) x = -x;
    }
  }
  std::sort(pairs.begin(), pairs.end(), [](const Pair &x, const Pair &y) {
    int64_t hx = x.a / 6, hy = y.a / 6, sx = x.b / 6, sy = y.b / 6;
    return std::tie(hx, sx, x.a) < std::tie(hy, sy, y.a);
  });

  auto make_face = [&](int64_t hslot, int64_t sslot, Vec shift, int32_t tag) {
    HexFace face;
    face.holder = hslot / 6;
    face.holder_face = int8_t(hslot % 6);
    face.sharer = sslot >= 0 ? sslot / 6 : -1;
    face.sharer_face = sslot >= 0 ? int8_t(sslot % 6) : -1;
    face.tag = tag;
    for (int d = 0; d < 3; ++d) face.shift[d] = shift[d];
    int fh = face.holder_face;
    // quadrature point order -> holder in-face coordinates
    int s0, t0, s1, t1;
    InFace(fh, kCorner[kFaceCorners[fh][0]], &s0, &t0);
    InFace(fh, kCorner[kFaceCorners[fh][1]], &s1, &t1);
    face.quad_to_holder = int8_t(OrientFrom(s0, t0, s1, t1));
    // frame of the holder-sorted face map at its centre
    Vec ra{ 0, 0, 0 }, rb{ 0, 0, 0 };
    for (int i = 0; i < 4; ++i) {
      Vec x = node(cell_node(face.holder, kFaceCorners[fh][i]));
      for (int d = 0; d < 3; ++d) {
        ra[d] += kQuadCorner[i][0] * x[d] / 4;
        rb[d] += kQuadCorner[i][1] * x[d] / 4;
      }
    }
    Vec n = Unit(Cross(ra, rb)), t = Unit(ra);
    for (int d = 0; d < 3; ++d) { face.frame[d] = n[d]; face.frame[3 + d] = t[d]; }
    for (int i = 0; i < 4; ++i) {
      Vec x = node(cell_node(face.holder, kFaceCorners[fh][i]));
      for (int d = 0; d < 3; ++d) face.corners[3 * i + d] = x[d];
    }
    face.orient = 0;
    if (face.sharer >= 0) {
      int fs = face.sharer_face;
      int img[2][2];
      const int probe[2][2] = { { -1, -1 }, { +1, -1 } };
      for (int k = 0; k < 2; ++k) {
        int ch = CornerAt(fh, probe[k][0], probe[k][1]);
        int64_t nid = cell_node(face.holder, ch);
        Vec pos = Add(node(nid), shift);
        int match = -1;
        for (int i = 0; i < 4; ++i) {
          int cs = kFaceCorners[fs][i];
          int64_t sid = cell_node(face.sharer, cs);
          bool same = (shift[0] == 0 && shift[1] == 0 && shift[2] == 0)
              ? sid == nid
              : Norm(Sub(node(sid), pos)) <= 1e-9 * (1 + Norm(pos));
          if (same) { match = cs; break; }
        }
        if (match < 0) throw std::invalid_argument("face corners do not match across a face");
        InFace(fs, kCorner[match], &img[k][0], &img[k][1]);
      }
      face.orient = int8_t(OrientFrom(img[0][0], img[0][1], img[1][0], img[1][1]));
    }
    return face;
  };
  out.faces.reserve(pairs.size() + bnd.size());
  for (auto &p : pairs) out.faces.push_back(make_face(p.a, p.b, p.shift, -1));
  out.n_interior = out.faces.size();
  for (auto &b : bnd)
    if (b.tag >= 0) out.faces.push_back(make_face(b.slot, -1, { 0, 0, 0 }, b.tag));
  return out;
}

}  // namespace mcfd
