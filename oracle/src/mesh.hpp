// Oracle (test infrastructure): mesh geometry and connectivity restated
// from mesh/part.hpp, coordinate/{hexahedron,quadrangle,tetrahedron,
// triangle}.hpp and integrator/{element,face,hexahedron,quadrangle}.hpp.
//
// Cell ids follow the input order (one rank: every cell is an inner cell,
// part.hpp:746-767).  Interior faces are built once per adjacency pair
// (m_holder < m_sharer) in the order "holder ascending, then sharer
// ascending" (part.hpp:932-975 consumes the pairs written by
// mesh/shuffler.hpp, which emits (i, j) with i < j while scanning cells i in
// order).  Boundary faces follow, in input order (part.hpp:1456-1549).
// Periodic boundary faces (tag < 0 on input) have no reference counterpart:
// each one is paired with the face on the opposite side of the box and the
// pair is treated as one more interior adjacency (holder = lower cell id).
#ifndef ORACLE_MESH_HPP_
#define ORACLE_MESH_HPP_

#include <cstdint>
#include <map>
#include <set>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "core.hpp"
#include "simplex.hpp"

namespace oracle {

enum CellType { kHexa8 = 0, kTetra4 = 1 };

// coordinate/hexahedron.hpp:353-359 (first 8 of Hexahedron27::local_coords_)
constexpr int kHexLocal[8][3] = {
  { -1, -1, -1 }, { +1, -1, -1 }, { +1, +1, -1 }, { -1, +1, -1 },
  { -1, -1, +1 }, { +1, -1, +1 }, { +1, +1, +1 }, { -1, +1, +1 },
};
// coordinate/hexahedron.hpp:379-389 (corner part of Hexahedron27::faces_)
constexpr int kHexFaces[6][4] = {
  { 0, 3, 2, 1 }, { 0, 1, 5, 4 }, { 1, 2, 6, 5 },
  { 2, 3, 7, 6 }, { 0, 4, 7, 3 }, { 4, 5, 6, 7 },
};
// coordinate/quadrangle.hpp:136-139
constexpr int kQuadLocal[4][2] = { { -1, -1 }, { +1, -1 }, { +1, +1 }, { -1, +1 } };

// coordinate/hexahedron.hpp:84-116 GetFaceId: sum of the cell-local indices
// of the first three cell nodes found on the face.
inline int HexGetFaceId(int64_t const *cell_nodes, int64_t const *face_nodes, int n_face_nodes) {
  int cnt = 0, c = 0, c_sum = 0;
  while (cnt < 3) {
    auto curr = cell_nodes[c];
    for (int f = 0; f < n_face_nodes; ++f) {
      if (face_nodes[f] == curr) { c_sum += c; ++cnt; break; }
    }
    ++c;
  }
  switch (c_sum) {
  case 3: return 0;
  case 5: return 1;
  case 7: return 4;
  case 8: return 2;
  case 11: return 3;
  case 15: return 5;
  default: throw std::runtime_error("HexGetFaceId: not a face");
  }
}

struct HexMap {  // coordinate/hexahedron.hpp:397-526 (Hexahedron8)
  Vec3 x[8];
  Vec3 LocalToGlobal(double a, double b, double c) const {
    Vec3 s{ 0, 0, 0 };
    for (int i = 0; i < 8; ++i) {
      double v = (1 + kHexLocal[i][0] * a);
      v *= (1 + kHexLocal[i][1] * b);
      v *= (1 + kHexLocal[i][2] * c);
      v /= 8;
      for (int d = 0; d < 3; ++d) s[d] += x[i][d] * v;
    }
    return s;
  }
  // rows are d/dxi, d/deta, d/dzeta of (x, y, z) (coordinate/cell.hpp:82-90)
  Mat3 LocalToJacobian(double a, double b, double c) const {
    Mat3 j{};
    for (int i = 0; i < 8; ++i) {
      double fx = 1 + kHexLocal[i][0] * a, fy = 1 + kHexLocal[i][1] * b, fz = 1 + kHexLocal[i][2] * c;
      double g[3] = { kHexLocal[i][0] * fy * fz / 8, kHexLocal[i][1] * fz * fx / 8, kHexLocal[i][2] * fx * fy / 8 };
      for (int r = 0; r < 3; ++r) for (int d = 0; d < 3; ++d) j[r][d] += g[r] * x[i][d];
    }
    return j;
  }
};

struct QuadMap {  // coordinate/quadrangle.hpp (Quadrangle4) + face.hpp:54-62
  Vec3 x[4];
  Vec3 LocalToGlobal(double a, double b) const {
    Vec3 s{ 0, 0, 0 };
    for (int i = 0; i < 4; ++i) {
      double v = (1 + kQuadLocal[i][0] * a) * (1 + kQuadLocal[i][1] * b) / 4;
      for (int d = 0; d < 3; ++d) s[d] += x[i][d] * v;
    }
    return s;
  }
  void Jacobian(double a, double b, Vec3 *ra, Vec3 *rb) const {
    *ra = { 0, 0, 0 }; *rb = { 0, 0, 0 };
    for (int i = 0; i < 4; ++i) {
      double ga = kQuadLocal[i][0] * (1 + kQuadLocal[i][1] * b) / 4;
      double gb = kQuadLocal[i][1] * (1 + kQuadLocal[i][0] * a) / 4;
      for (int d = 0; d < 3; ++d) { (*ra)[d] += ga * x[i][d]; (*rb)[d] += gb * x[i][d]; }
    }
  }
  Frame NormalFrame(double a, double b) const {
    Vec3 ra, rb;
    Jacobian(a, b, &ra, &rb);
    Frame f;
    f.a = normalized(cross(ra, rb));
    f.b = normalized(ra);
    f.c = cross(f.a, f.b);
    return f;
  }
  double AreaFactor(double a, double b) const {  // sqrt(det(J J^T)), integrator/element.hpp:108-110
    Vec3 ra, rb;
    Jacobian(a, b, &ra, &rb);
    double g11 = dot(ra, ra), g12 = dot(ra, rb), g22 = dot(rb, rb);
    return std::sqrt(g11 * g22 - g12 * g12);
  }
};

// coordinate/tetrahedron.hpp:136-148: local nodes (1,0,0), (0,1,0), (0,0,1),
// (0,0,0); faces listed so that the face map's normal points out of a
// CGNS-ordered TETRA_4.
constexpr int kTetFaces[4][3] = { { 0, 2, 1 }, { 0, 1, 3 }, { 0, 3, 2 }, { 1, 2, 3 } };

// coordinate/tetrahedron.hpp:77-97 (SortNodesOnFace): the face id is the
// sum of the cell-local indices of its three nodes minus 3.
inline int TetGetFaceId(int64_t const *cell_nodes, int64_t const *face_nodes) {
  int sum = 0, cnt = 0;
  for (int c = 0; c < 4 && cnt < 3; ++c)
    for (int f = 0; f < 3; ++f)
      if (face_nodes[f] == cell_nodes[c]) { sum += c; ++cnt; break; }
  if (cnt != 3) throw std::runtime_error("TetGetFaceId: not a face");
  return sum - 3;
}

struct TetMap {  // coordinate/tetrahedron.hpp:56-130 (Tetrahedron4), cell.hpp:82-90
  Vec3 x[4];
  Vec3 LocalToGlobal(double a, double b, double c) const {
    const double sh[4] = { a, b, c, 1.0 - a - b - c };
    Vec3 s = scale(x[0], sh[0]);
    for (int i = 1; i < 4; ++i) s = add(s, scale(x[i], sh[i]));
    return s;
  }
  Mat3 Jacobian() const {  // rows d/dxi, d/deta, d/dzeta
    Mat3 j;
    for (int r = 0; r < 3; ++r) j[r] = sub(x[r], x[3]);
    return j;
  }
};

struct TriMap {  // coordinate/triangle.hpp (Triangle3) + face.hpp:54-62
  Vec3 x[3];
  Vec3 LocalToGlobal(double a, double b) const {
    Vec3 s = scale(x[0], a);
    s = add(s, scale(x[1], b));
    return add(s, scale(x[2], 1.0 - a - b));
  }
  Frame NormalFrame() const {
    Vec3 ra = sub(x[0], x[2]), rb = sub(x[1], x[2]);
    Frame f;
    f.a = normalized(cross(ra, rb));
    f.b = normalized(ra);
    f.c = cross(f.a, f.b);
    return f;
  }
  double AreaFactor() const {
    Vec3 ra = sub(x[0], x[2]), rb = sub(x[1], x[2]);
    double g11 = dot(ra, ra), g12 = dot(ra, rb), g22 = dot(rb, rb);
    return std::sqrt(g11 * g22 - g12 * g12);
  }
};

struct CellGeom {
  int type;
  int64_t nodes[8];
  HexMap map;              // hexahedra
  TetMap tmap;             // tetrahedra
  Vec3 center;
  double volume, length;   // length = cbrt(volume) / 2 (part.hpp:230)
  std::vector<Vec3> qp;    // quadrature points (global)
  std::vector<double> qw;  // global weights
  std::vector<double> qw_local;
};

struct FaceGeom {
  int64_t holder, sharer;  // sharer = -1 on boundaries
  int holder_face, sharer_face;
  int tag;                 // boundary tag (interior faces: -1)
  int64_t nodes[4];        // sorted w.r.t. the holder (SortNodesOnFace)
  Vec3 shift;              // sharer-side coordinates = holder-side + shift
  QuadMap map;             // quadrangles (hexahedral meshes)
  TriMap tmap;             // triangles (tetrahedral meshes)
  std::vector<Vec3> qp;
  std::vector<double> qw;
  std::vector<Frame> frames;
  Vec3 center;
};

struct Mesh {
  int cell_type = kHexa8;
  int q = 0;                 // 1-D quadrature size (hexahedra)
  bool lobatto = false;
  std::vector<Vec3> coords;
  std::vector<CellGeom> cells;
  std::vector<FaceGeom> faces;  // interior first, then boundary
  int64_t n_interior = 0;
  double x1d[8], w1d[8];
};

struct MeshInput {
  int cell_type;
  int64_t n_nodes;
  double const *coords;
  int64_t n_cells;
  int64_t const *cells;
  int64_t n_bfaces;
  int64_t const *bface_nodes;
  int32_t const *bface_tags;  // < 0: periodic
  double const *period;       // [3], 0 for a non-periodic direction
};

inline void BuildHexCellGeometry(Mesh *mesh, CellGeom *cell) {
  int q = mesh->q;
  for (int i = 0; i < 8; ++i) cell->map.x[i] = mesh->coords[cell->nodes[i]];
  cell->center = cell->map.LocalToGlobal(0, 0, 0);
  cell->qp.clear(); cell->qw.clear(); cell->qw_local.clear();
  double vol = 0;
  // integrator/hexahedron.hpp:66-93 (point order) + element.hpp:102-115
  for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) for (int k = 0; k < q; ++k) {
    double a = mesh->x1d[i], b = mesh->x1d[j], c = mesh->x1d[k];
    double wl = mesh->w1d[i] * mesh->w1d[j] * mesh->w1d[k];
    cell->qp.push_back(cell->map.LocalToGlobal(a, b, c));
    double wg = wl * det3(cell->map.LocalToJacobian(a, b, c));
    cell->qw.push_back(wg);
    cell->qw_local.push_back(wl);
    vol += wg;
  }
  cell->volume = vol;
  cell->length = std::cbrt(vol) * 0.5;
}

inline void BuildQuadFaceGeometry(Mesh *mesh, FaceGeom *face) {
  int q = mesh->q;
  for (int i = 0; i < 4; ++i) face->map.x[i] = mesh->coords[face->nodes[i]];
  face->center = face->map.LocalToGlobal(0, 0);
  // integrator/quadrangle.hpp:62-83 (point order)
  for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) {
    double a = mesh->x1d[i], b = mesh->x1d[j];
    face->qp.push_back(face->map.LocalToGlobal(a, b));
    face->qw.push_back(mesh->w1d[i] * mesh->w1d[j] * face->map.AreaFactor(a, b));
    face->frames.push_back(face->map.NormalFrame(a, b));
  }
}

// integrator/tetrahedron.hpp:89-92 + element.hpp:102-115 with the 14-point
// rule.  The weight is |det J| x local weight: the reference element's node
// order is left-handed with respect to CGNS TETRA_4, so the signed
// determinant is negative for a CGNS-ordered cell, while the reference's own
// test (test/integrator/tetrahedron.cpp:25) expects the positive volume.
inline void BuildTetCellGeometry(Mesh *mesh, CellGeom *cell) {
  static const TetRule14 rule;
  for (int i = 0; i < 4; ++i) cell->tmap.x[i] = mesh->coords[cell->nodes[i]];
  cell->center = cell->tmap.LocalToGlobal(0.25, 0.25, 0.25);
  cell->qp.clear(); cell->qw.clear(); cell->qw_local.clear();
  const double det = std::abs(det3(cell->tmap.Jacobian()));
  double vol = 0;
  for (int q = 0; q < 14; ++q) {
    cell->qp.push_back(cell->tmap.LocalToGlobal(rule.x[q][0], rule.x[q][1], rule.x[q][2]));
    double wg = rule.w[q] * det;
    cell->qw.push_back(wg);
    cell->qw_local.push_back(rule.w[q]);
    vol += wg;
  }
  cell->volume = vol;
  cell->length = std::cbrt(vol) * 0.5;
}

inline void BuildTriFaceGeometry(Mesh *mesh, FaceGeom *face) {
  static const TriRule6 rule;
  for (int i = 0; i < 3; ++i) face->tmap.x[i] = mesh->coords[face->nodes[i]];
  face->center = face->tmap.LocalToGlobal(1.0 / 3, 1.0 / 3);
  const double area = face->tmap.AreaFactor();
  const Frame frame = face->tmap.NormalFrame();
  for (int q = 0; q < 6; ++q) {
    face->qp.push_back(face->tmap.LocalToGlobal(rule.x[q][0], rule.x[q][1]));
    face->qw.push_back(rule.w[q] * area);
    face->frames.push_back(frame);
  }
}

// Build the mesh: cells, interior faces (adjacency order), boundary faces.
// Hexahedra use q x q x q Gauss points (Lobatto if asked); tetrahedra the
// 14-point rule with 6-point triangles on their faces.
inline Mesh BuildMesh(MeshInput const &in, int q, bool lobatto) {
  Mesh mesh;
  const bool tet = in.cell_type == kTetra4;
  const int npe = tet ? 4 : 8, fnpe = tet ? 3 : 4;
  mesh.cell_type = tet ? kTetra4 : kHexa8;
  mesh.q = q;
  mesh.lobatto = lobatto;
  if (lobatto) LobattoRule(q, mesh.x1d, mesh.w1d); else LegendreRule(q, mesh.x1d, mesh.w1d);
  mesh.coords.resize(in.n_nodes);
  for (int64_t i = 0; i < in.n_nodes; ++i)
    mesh.coords[i] = { in.coords[3 * i], in.coords[3 * i + 1], in.coords[3 * i + 2] };
  mesh.cells.resize(in.n_cells);
  std::unordered_map<int64_t, std::vector<int64_t>> node_to_cells;
  for (int64_t c = 0; c < in.n_cells; ++c) {
    auto &cell = mesh.cells[c];
    cell.type = mesh.cell_type;
    for (int i = 0; i < npe; ++i) {
      cell.nodes[i] = in.cells[npe * c + i];
      node_to_cells[cell.nodes[i]].push_back(c);
    }
    if (tet) BuildTetCellGeometry(&mesh, &cell); else BuildHexCellGeometry(&mesh, &cell);
  }
  // Adjacency by >= 3 common nodes (mesh/shuffler.hpp:526-527), one entry
  // per pair (i < j) with the holder's local face id.
  std::map<std::pair<int64_t, int64_t>, std::vector<std::tuple<int, int, Vec3>>> adj;
  for (int64_t c = 0; c < in.n_cells; ++c) {
    std::map<int64_t, int> cnt;
    for (int i = 0; i < npe; ++i)
      for (auto other : node_to_cells[mesh.cells[c].nodes[i]])
        if (other > c) cnt[other]++;
    for (auto [other, n] : cnt) {
      if (n < 3) continue;
      std::vector<int64_t> common;
      for (int i = 0; i < npe; ++i)
        for (int k = 0; k < npe; ++k)
          if (mesh.cells[c].nodes[i] == mesh.cells[other].nodes[k]) common.push_back(mesh.cells[c].nodes[i]);
      int fh = tet ? TetGetFaceId(mesh.cells[c].nodes, common.data())
                   : HexGetFaceId(mesh.cells[c].nodes, common.data(), (int)common.size());
      int fs = tet ? TetGetFaceId(mesh.cells[other].nodes, common.data())
                   : HexGetFaceId(mesh.cells[other].nodes, common.data(), (int)common.size());
      adj[{ c, other }].emplace_back(fh, fs, Vec3{ 0, 0, 0 });
    }
  }
  // Boundary faces: holder = the cell holding all face nodes.
  struct BFace { int64_t holder; int face; int tag; int64_t nodes[4]; Vec3 center; };
  std::vector<BFace> bfaces;
  for (int64_t b = 0; b < in.n_bfaces; ++b) {
    BFace bf;
    for (int i = 0; i < fnpe; ++i) bf.nodes[i] = in.bface_nodes[fnpe * b + i];
    bf.tag = in.bface_tags[b];
    if (tet && bf.tag < 0) throw std::runtime_error("periodic faces are supported on hexahedra only");
    std::map<int64_t, int> cnt;
    for (int i = 0; i < fnpe; ++i) for (auto c : node_to_cells[bf.nodes[i]]) cnt[c]++;
    bf.holder = -1;
    for (auto [c, n] : cnt) if (n == fnpe) { bf.holder = c; break; }
    if (bf.holder < 0) throw std::runtime_error("boundary face without holder");
    bf.face = tet ? TetGetFaceId(mesh.cells[bf.holder].nodes, bf.nodes)
                  : HexGetFaceId(mesh.cells[bf.holder].nodes, bf.nodes, 4);
    bf.center = { 0, 0, 0 };
    for (int i = 0; i < fnpe; ++i) bf.center = add(bf.center, scale(mesh.coords[bf.nodes[i]], 1.0 / fnpe));
    bfaces.push_back(bf);
  }
  // Periodic pairing: translate the center of each periodic face by +L_d.
  std::vector<int> paired(bfaces.size(), 0);
  for (size_t a = 0; a < bfaces.size(); ++a) {
    if (bfaces[a].tag >= 0 || paired[a]) continue;
    for (size_t b = a + 1; b < bfaces.size(); ++b) {
      if (bfaces[b].tag >= 0 || paired[b]) continue;
      for (int d = 0; d < 3; ++d) {
        if (!(in.period[d] > 0)) continue;
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          Vec3 s{ 0, 0, 0 };
          s[d] = sgn * in.period[d];
          if (norm(sub(add(bfaces[a].center, s), bfaces[b].center)) < 1e-10 * (1 + in.period[d])) {
            paired[a] = paired[b] = 1;
            int64_t ca = bfaces[a].holder, cb = bfaces[b].holder;
            if (ca < cb) adj[{ ca, cb }].emplace_back(bfaces[a].face, bfaces[b].face, s);
            else adj[{ cb, ca }].emplace_back(bfaces[b].face, bfaces[a].face, scale(s, -1));
            goto next_face;
          }
        }
      }
    }
    throw std::runtime_error("unpaired periodic face");
  next_face:;
  }
  for (auto &[pair, list] : adj) {
    auto sorted = list;
    std::sort(sorted.begin(), sorted.end(), [](auto const &x, auto const &y) {
      return std::get<0>(x) < std::get<0>(y);
    });
    for (auto &[fh, fs, s] : sorted) {
      FaceGeom face;
      face.holder = pair.first; face.sharer = pair.second;
      face.holder_face = fh; face.sharer_face = fs; face.tag = -1; face.shift = s;
      for (int i = 0; i < fnpe; ++i)
        face.nodes[i] = mesh.cells[pair.first].nodes[tet ? kTetFaces[fh][i] : kHexFaces[fh][i]];
      if (tet) BuildTriFaceGeometry(&mesh, &face); else BuildQuadFaceGeometry(&mesh, &face);
      mesh.faces.push_back(std::move(face));
    }
  }
  mesh.n_interior = mesh.faces.size();
  for (size_t b = 0; b < bfaces.size(); ++b) {
    if (bfaces[b].tag < 0) continue;
    FaceGeom face;
    face.holder = bfaces[b].holder; face.sharer = -1;
    face.holder_face = bfaces[b].face; face.sharer_face = -1; face.tag = bfaces[b].tag;
    face.shift = { 0, 0, 0 };
    for (int i = 0; i < fnpe; ++i)
      face.nodes[i] = mesh.cells[face.holder].nodes[tet ? kTetFaces[face.holder_face][i]
                                                        : kHexFaces[face.holder_face][i]];
    if (tet) BuildTriFaceGeometry(&mesh, &face); else BuildQuadFaceGeometry(&mesh, &face);
    mesh.faces.push_back(std::move(face));
  }
  return mesh;
}

inline Mesh BuildHexMesh(MeshInput const &in, int q, bool lobatto) {
  if (in.cell_type != kHexa8) throw std::runtime_error("nodal FR needs hexahedra");
  return BuildMesh(in, q, lobatto);
}

}  // namespace oracle

#endif  // ORACLE_MESH_HPP_
