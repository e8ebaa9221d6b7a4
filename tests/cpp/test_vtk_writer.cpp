// mini::mesh::vtk::Writer on a host-only FiniteElement: reads a mesh and a
// solution column from the file named by argv[1] (int64 n_nodes, n_cells,
// n_bfaces, degree, scheme; then coords, cell nodes, boundary face nodes,
// boundary tags as int64, and the column), writes one frame with a pressure
// point field and a cell-index cell field into directory argv[2] as "frame".
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <vector>

#include "mini/mesh/vtk.hpp"

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  std::ifstream in(argv[1], std::ios::binary);
  int64_t head[5];
  in.read(reinterpret_cast<char *>(head), sizeof(head));
  const int64_t n_nodes = head[0], n_cells = head[1], n_bfaces = head[2];
  const int degree = (int)head[3];
  const bool tet = head[4] != 0;
  const int npe = tet ? 4 : 8, fpe = tet ? 3 : 4;
  std::vector<double> coords(n_nodes * 3);
  std::vector<int64_t> cells(n_cells * npe), bnodes(n_bfaces * fpe), btags64(n_bfaces);
  in.read(reinterpret_cast<char *>(coords.data()), coords.size() * 8);
  in.read(reinterpret_cast<char *>(cells.data()), cells.size() * 8);
  in.read(reinterpret_cast<char *>(bnodes.data()), bnodes.size() * 8);
  in.read(reinterpret_cast<char *>(btags64.data()), btags64.size() * 8);
  std::vector<int32_t> btags(btags64.begin(), btags64.end());
  mcfd_mesh_desc mesh{};
  mesh.cell_type = tet ? MCFD_TETRA_4 : MCFD_HEXA_8;
  mesh.n_nodes = n_nodes; mesh.coords = coords.data();
  mesh.n_cells = n_cells; mesh.cell_nodes = cells.data();
  mesh.n_bfaces = n_bfaces; mesh.bface_nodes = bnodes.data(); mesh.bface_tags = btags.data();
  mcfd_scheme_desc scheme{};
  scheme.scheme = tet ? MCFD_SCHEME_DG_GENERAL : MCFD_SCHEME_FR_GENERAL;
  scheme.degree = degree;
  scheme.nodes = MCFD_NODES_LEGENDRE;
  scheme.riemann = MCFD_RIEMANN_HLLC;
  scheme.precision = MCFD_PRECISION_FP64;
  scheme.c_next = 1.0;
  scheme.flags = MCFD_FLAG_HOST_ONLY;
  scheme.weno_w0 = 0.001;
  scheme.weno_eps = 1e-6;
  std::vector<mcfd_bc_desc> bcs(6, mini::spatial::mi355x::InviscidWall());
  try {
    mini::spatial::mi355x::FiniteElement fe(mesh, scheme, bcs);
    std::vector<double> column(fe.ColumnSize());
    in.read(reinterpret_cast<char *>(column.data()), column.size() * 8);
    if (!in) {
      std::printf("FAIL short input\n");
      return 1;
    }
    using Writer = mini::mesh::vtk::Writer<>;
    Writer::AddPointData("Pressure", [](double const *, double const *v) {
      return 0.4 * (v[4] - 0.5 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) / v[0]);
    });
    Writer::AddCellData("CellId", [](int64_t c) { return double(c); });
    std::printf("PASS wrote %s\n", Writer::WriteSolutions(fe, argv[2], "frame", column.data()).c_str());
  } catch (std::exception const &e) {
    std::printf("FAIL %s\n", e.what());
    return 1;
  }
  return 0;
}
