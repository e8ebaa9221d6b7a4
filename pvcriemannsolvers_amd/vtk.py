"""Frame output: the solution as VTK XML unstructured grids.

Mirrors ``Writer<Part>::WriteSolutions`` (mesh/vtk.hpp:375-681): every cell
becomes one second-order VTK cell whose nodes carry the solution evaluated
there (``polynomial().LocalToGlobalAndValue``, vtk.hpp:431-460):

* hexahedra -> VTK_TRIQUADRATIC_HEXAHEDRON (type 29, 27 nodes, the local
  points of vtk.hpp:310-330);
* tetrahedra -> VTK_QUADRATIC_TETRA (type 24, 10 nodes, vtk.hpp:209-221).

Files, as the reference lays them out for one rank:
``{directory}/{soln_name}.pvtu`` names the piece ``{directory}/{soln_name}/0.vtu``.
Point data are the five conservative components under the names the Euler
demos give them (demo/euler/sourceless.cpp:109-110), then any extra point
fields; cell data are the extra cell fields.  Points are written as base64
binary with a UInt64 byte-count header, the rest as ASCII (vtk.hpp:589-664);
ASCII values are written with 17 significant digits, so a frame round-trips
the FP64 values (the reference streams them at the stream's default
precision).
"""
from __future__ import annotations

import base64
import os
import sys

import numpy as np

FIELD_NAMES = ("Density", "MomentumX", "MomentumY", "MomentumZ", "EnergyStagnationDensity")

# vtk.hpp:310-330 (VTK node order of vtkTriQuadraticHexahedron)
HEX27_LOCAL = np.array([
    (-1, -1, -1), (+1, -1, -1), (+1, +1, -1), (-1, +1, -1),
    (-1, -1, +1), (+1, -1, +1), (+1, +1, +1), (-1, +1, +1),
    (0, -1, -1), (+1, 0, -1), (0, +1, -1), (-1, 0, -1),
    (0, -1, +1), (+1, 0, +1), (0, +1, +1), (-1, 0, +1),
    (-1, -1, 0), (+1, -1, 0), (+1, +1, 0), (-1, +1, 0),
    (-1, 0, 0), (+1, 0, 0), (0, -1, 0), (0, +1, 0), (0, 0, -1), (0, 0, +1),
    (0, 0, 0),
], dtype=np.float64)
# vtk.hpp:209-221 (VTK node order of vtkQuadraticTetra)
TET10_LOCAL = np.array([
    (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0),
    (0.5, 0.5, 0), (0, 0.5, 0.5), (0.5, 0, 0.5),
    (0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5),
], dtype=np.float64)
VTK_QUADRATIC_TETRA = 24
VTK_TRIQUADRATIC_HEXAHEDRON = 29


def _ascii(values) -> str:
    return " ".join(repr(float(v)) for v in np.asarray(values).ravel())


def _binary(array: np.ndarray) -> str:
    """Base64 of the UInt64 byte count, then base64 of the bytes (vtk.hpp:631-640)."""
    data = np.ascontiguousarray(array, dtype="<f8").tobytes()
    header = base64.b64encode(np.array([len(data)], dtype="<u8").tobytes()).decode()
    return header + base64.b64encode(data).decode()


def write_solutions(fe, directory: str, soln_name: str, point_data=(), cell_data=(),
                    field_names=FIELD_NAMES, column=None) -> str:
    """Writes one frame of `fe` (a spatial.FiniteElement) and returns the path of the .pvtu.

    point_data: (name, f) pairs, f(xyz [n, 3], values [n, 5]) -> [n], the
    Writer::AddPointData fields; cell_data: (name, f) pairs, f(cell ids [n]) -> [n],
    the Writer::AddCellData fields.  `column` replaces the handle's current
    solution (a host-only handle needs one).
    """
    hexa = fe.cell_type == 0
    local = HEX27_LOCAL if hexa else TET10_LOCAL
    vtk_type = VTK_TRIQUADRATIC_HEXAHEDRON if hexa else VTK_QUADRATIC_TETRA
    xyz, values = fe.evaluate(local, column)
    n_cells, n_local = xyz.shape[0], xyz.shape[1]
    xyz, values = xyz.reshape(-1, 3), values.reshape(-1, 5)
    little = sys.byteorder == "little"
    order = '"LittleEndian"' if little else '"BigEndian"'
    os.makedirs(os.path.join(directory, soln_name), exist_ok=True)
    pvtu = os.path.join(directory, soln_name + ".pvtu")
    with open(pvtu, "w") as f:
        f.write(f'<VTKFile type="PUnstructuredGrid" version="1.0" byte_order={order} header_type="UInt64">\n')
        f.write('  <PUnstructuredGrid GhostLevel="1">\n    <PPointData>\n')
        for name in list(field_names) + [n for n, _ in point_data]:
            f.write(f'      <PDataArray type="Float64" Name="{name}"/>\n')
        f.write("    </PPointData>\n    <PCellData>\n")
        for name, _ in cell_data:
            f.write(f'      <PDataArray type="Float64" Name="{name}"/>\n')
        f.write("    </PCellData>\n    <PPoints>\n")
        f.write('      <PDataArray type="Float64" Name="Points" NumberOfComponents="3"/>\n    </PPoints>\n')
        f.write(f'    <Piece Source="./{soln_name}/0.vtu"/>\n  </PUnstructuredGrid>\n</VTKFile>\n')
    fmt = '"ascii"'
    with open(os.path.join(directory, soln_name, "0.vtu"), "w") as f:
        f.write(f'<VTKFile type="UnstructuredGrid" version="1.0" byte_order={order} header_type="UInt64">\n')
        f.write(f'  <UnstructuredGrid>\n    <Piece NumberOfPoints="{len(xyz)}" NumberOfCells="{n_cells}">\n')
        f.write("      <PointData>\n")
        for k, name in enumerate(field_names):
            f.write(f'        <DataArray type="Float64" Name="{name}" format={fmt}>\n')
            f.write(_ascii(values[:, k]) + "\n        </DataArray>\n")
        for name, fn in point_data:
            f.write(f'        <DataArray type="Float64" Name="{name}" format={fmt}>\n')
            f.write(_ascii(fn(xyz, values)) + "\n        </DataArray>\n")
        f.write("      </PointData>\n      <CellData>\n")
        for name, fn in cell_data:
            f.write(f'        <DataArray type="Float64" Name="{name}" format={fmt}>\n')
            f.write(_ascii(fn(np.arange(n_cells))) + "\n        </DataArray>\n")
        f.write("      </CellData>\n      <Points>\n")
        f.write('        <DataArray type="Float64" Name="Points" NumberOfComponents="3" format="binary">\n')
        f.write(_binary(xyz) + "\n        </DataArray>\n      </Points>\n      <Cells>\n")
        f.write('        <DataArray type="Int32" Name="connectivity" format="ascii">\n')
        f.write(" ".join(map(str, range(len(xyz)))) + "\n        </DataArray>\n")
        f.write(f'        <DataArray type="Int32" Name="offsets" format={fmt}>\n')
        f.write(" ".join(str(n_local * (c + 1)) for c in range(n_cells)) + "\n        </DataArray>\n")
        f.write(f'        <DataArray type="UInt8" Name="types" format={fmt}>\n')
        f.write(" ".join([str(vtk_type)] * n_cells) + "\n        </DataArray>\n")
        f.write("      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")
    return pvtu


def read_piece(path: str) -> dict:
    """Parses a .vtu piece written by write_solutions: points, point and cell
    fields by name, connectivity, offsets and types."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {"point_data": {}, "cell_data": {}}
    for arr in piece.find("PointData"):
        out["point_data"][arr.get("Name")] = np.array(arr.text.split(), dtype=np.float64)
    for arr in piece.find("CellData"):
        out["cell_data"][arr.get("Name")] = np.array(arr.text.split(), dtype=np.float64)
    pts = piece.find("Points/DataArray").text.strip()
    n_bytes = int(np.frombuffer(base64.b64decode(pts[:12]), dtype="<u8")[0])
    data = base64.b64decode(pts[12:])
    assert len(data) == n_bytes
    out["points"] = np.frombuffer(data, dtype="<f8").reshape(-1, 3)
    for arr in piece.find("Cells"):
        out[arr.get("Name")] = np.array(arr.text.split(), dtype=np.int64)
    return out
