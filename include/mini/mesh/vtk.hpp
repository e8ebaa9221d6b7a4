// mini::mesh::vtk::Writer<Spatial> for a mini::spatial::mi355x::FiniteElement: frame
// output as VTK XML unstructured grids, in the layout of the reference's
// Writer<Part>::WriteSolutions (mesh/vtk.hpp:375-681) for one rank:
// `{directory}/{soln_name}.pvtu` naming the piece `{directory}/{soln_name}/0.vtu`.
//
// Every cell becomes one second-order VTK cell whose nodes carry the solution
// evaluated there (FiniteElement::Evaluate, polynomial().LocalToGlobalAndValue):
// hexahedra as VTK_TRIQUADRATIC_HEXAHEDRON (27 nodes, the local points of
// mesh/vtk.hpp:310-330), tetrahedra as VTK_QUADRATIC_TETRA (10 nodes,
// vtk.hpp:209-221).  Point data are the five components under the element's
// field names, then the fields added by AddPointData; cell data the fields
// added by AddCellData.  Points are base64 binary with a UInt64 byte-count
// header; values are ASCII with 17 significant digits.  A point field is
// f(xyz, value) of a frame node and a cell field f(cell index), in place of
// the reference's f(cell, xyz, value) and f(cell).
#ifndef MINI_MESH_VTK_HPP_
#define MINI_MESH_VTK_HPP_

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mini/spatial/mi355x.hpp"

namespace mini {
namespace mesh {
namespace vtk {

// RFC 4648 base64 of n bytes.
inline std::string EncodeBase64(const unsigned char *data, std::size_t n) {
  static const char kTable[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  std::size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const unsigned v = (unsigned)data[i] << 16 | (unsigned)data[i + 1] << 8 | data[i + 2];
    out += kTable[v >> 18 & 63]; out += kTable[v >> 12 & 63]; out += kTable[v >> 6 & 63]; out += kTable[v & 63];
  }
  if (i < n) {
    const unsigned v = (unsigned)data[i] << 16 | (i + 1 < n ? (unsigned)data[i + 1] << 8 : 0u);
    out += kTable[v >> 18 & 63];
    out += kTable[v >> 12 & 63];
    out += i + 1 < n ? kTable[v >> 6 & 63] : '=';
    out += '=';
  }
  return out;
}

template <class Spatial = spatial::mi355x::FiniteElement>
class Writer {
 public:
  // f(xyz[3], value[5]) at a node of the frame (Writer::AddPointData, vtk.hpp:494-506)
  using PointData = std::pair<std::string, std::function<double(double const *, double const *)>>;
  // f(cell index) (Writer::AddCellData, vtk.hpp:508-520)
  using CellData = std::pair<std::string, std::function<double(int64_t)>>;

  template <class F>
  static void AddPointData(std::string const &name, F &&f) {
    point_data().emplace_back(name, std::forward<F>(f));
  }
  template <class F>
  static void AddCellData(std::string const &name, F &&f) {
    cell_data().emplace_back(name, std::forward<F>(f));
  }

  // The local points of one frame cell: hexahedra (vtk.hpp:310-330) or tetrahedra (vtk.hpp:209-221).
  static std::vector<double> LocalPoints(int32_t cell_type) {
    if (cell_type == MCFD_HEXA_8) {
      return { -1, -1, -1, +1, -1, -1, +1, +1, -1, -1, +1, -1, -1, -1, +1, +1, -1, +1, +1, +1, +1, -1, +1, +1,
               0, -1, -1, +1, 0, -1, 0, +1, -1, -1, 0, -1, 0, -1, +1, +1, 0, +1, 0, +1, +1, -1, 0, +1,
               -1, -1, 0, +1, -1, 0, +1, +1, 0, -1, +1, 0,
               -1, 0, 0, +1, 0, 0, 0, -1, 0, 0, +1, 0, 0, 0, -1, 0, 0, +1,
               0, 0, 0 };
    }
    return { 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0,
             0.5, 0.5, 0, 0, 0.5, 0.5, 0.5, 0, 0.5, 0.5, 0, 0, 0, 0.5, 0, 0, 0, 0.5 };
  }

  // Writes one frame of `spatial` (its current solution, or `column`);
  // returns the path of the .pvtu file.
  static std::string WriteSolutions(Spatial const &spatial, std::string const &directory,
                                    std::string const &soln_name, double const *column = nullptr) {
    const int32_t type = spatial.CellType();
    const std::vector<double> local = LocalPoints(type);
    const int n_local = (int)(local.size() / 3);
    const int vtk_type = type == MCFD_HEXA_8 ? 29 : 24;
    std::vector<double> xyz, values;
    spatial.Evaluate(local, &xyz, &values, column);
    const int64_t n_cells = spatial.CountCells(), n_points = n_cells * n_local;
    std::filesystem::create_directories(std::filesystem::path(directory) / soln_name);
    const std::string order = "\"LittleEndian\"";
    const std::string pvtu = directory + "/" + soln_name + ".pvtu";
    {
      std::ofstream f(pvtu);
      f << "<VTKFile type=\"PUnstructuredGrid\" version=\"1.0\" byte_order=" << order
        << " header_type=\"UInt64\">\n  <PUnstructuredGrid GhostLevel=\"1\">\n    <PPointData>\n";
      for (int k = 0; k < 5; ++k)
        f << "      <PDataArray type=\"Float64\" Name=\"" << spatial.GetFieldName(k) << "\"/>\n";
      for (auto &[name, fn] : point_data()) f << "      <PDataArray type=\"Float64\" Name=\"" << name << "\"/>\n";
      f << "    </PPointData>\n    <PCellData>\n";
      for (auto &[name, fn] : cell_data()) f << "      <PDataArray type=\"Float64\" Name=\"" << name << "\"/>\n";
      f << "    </PCellData>\n    <PPoints>\n"
        << "      <PDataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>\n    </PPoints>\n"
        << "    <Piece Source=\"./" << soln_name << "/0.vtu\"/>\n  </PUnstructuredGrid>\n</VTKFile>\n";
      if (!f) throw std::runtime_error("cannot write " + pvtu);
    }
    const std::string vtu = directory + "/" + soln_name + "/0.vtu";
    std::ofstream f(vtu);
    char buf[32];
    auto put = [&](double v) {
      std::snprintf(buf, sizeof(buf), "%.17g ", v);
      f << buf;
    };
    f << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=" << order
      << " header_type=\"UInt64\">\n  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"" << n_points
      << "\" NumberOfCells=\"" << n_cells << "\">\n      <PointData>\n";
    for (int k = 0; k < 5; ++k) {
      f << "        <DataArray type=\"Float64\" Name=\"" << spatial.GetFieldName(k) << "\" format=\"ascii\">\n";
      for (int64_t p = 0; p < n_points; ++p) put(values[p * 5 + k]);
      f << "\n        </DataArray>\n";
    }
    for (auto &[name, fn] : point_data()) {
      f << "        <DataArray type=\"Float64\" Name=\"" << name << "\" format=\"ascii\">\n";
      for (int64_t p = 0; p < n_points; ++p) put(fn(&xyz[p * 3], &values[p * 5]));
      f << "\n        </DataArray>\n";
    }
    f << "      </PointData>\n      <CellData>\n";
    for (auto &[name, fn] : cell_data()) {
      f << "        <DataArray type=\"Float64\" Name=\"" << name << "\" format=\"ascii\">\n";
      for (int64_t c = 0; c < n_cells; ++c) put(fn(c));
      f << "\n        </DataArray>\n";
    }
    f << "      </CellData>\n      <Points>\n"
      << "        <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"binary\">\n";
    // vtk.hpp:631-640: base64 of the byte count, then base64 of the bytes
    const uint64_t n_bytes = (uint64_t)xyz.size() * sizeof(double);
    unsigned char header[8];
    std::memcpy(header, &n_bytes, 8);
    f << EncodeBase64(header, 8)
      << EncodeBase64(reinterpret_cast<const unsigned char *>(xyz.data()), (std::size_t)n_bytes);
    f << "\n        </DataArray>\n      </Points>\n      <Cells>\n"
      << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int64_t p = 0; p < n_points; ++p) f << p << ' ';
    f << "\n        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (int64_t c = 0; c < n_cells; ++c) f << n_local * (c + 1) << ' ';
    f << "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (int64_t c = 0; c < n_cells; ++c) f << vtk_type << ' ';
    f << "\n        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    if (!f) throw std::runtime_error("cannot write " + vtu);
    return pvtu;
  }

 private:
  static std::vector<PointData> &point_data() {
    static std::vector<PointData> v;
    return v;
  }
  static std::vector<CellData> &cell_data() {
    static std::vector<CellData> v;
    return v;
  }
};

}  // namespace vtk
}  // namespace mesh
}  // namespace mini

#endif  // MINI_MESH_VTK_HPP_
