"""mini::spatial surface backed by libminicfd.

``FiniteElement`` implements :class:`temporal.System` exactly like the
reference's ``FiniteElement`` (spatial/fem.hpp:224-258): columns cross the
boundary by value.  ``rk3_update`` is the fused device path of
``RungeKutta<3>::Update`` (temporal/rk.hpp:31-49) that keeps the state in HBM.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C
from .temporal import System


def _mesh_desc(mesh: dict, keep: list) -> "C.MeshDesc":
    arrs = [np.ascontiguousarray(mesh["coords"], np.float64), np.ascontiguousarray(mesh["cells"], np.int64),
            np.ascontiguousarray(mesh["bface_nodes"], np.int64), np.ascontiguousarray(mesh["bface_tags"], np.int32)]
    keep.extend(arrs)
    coords, cells, bn, bt = arrs
    md = C.MeshDesc()
    md.cell_type = int(mesh.get("cell_type", C.HEXA_8))
    md.n_nodes, md.coords = len(coords), C.ptr(coords, C._dp)
    md.n_cells, md.cell_nodes = len(cells), C.ptr(cells, C._i64p)
    md.n_bfaces, md.bface_nodes, md.bface_tags = len(bn), C.ptr(bn, C._i64p), C.ptr(bt, C._i32p)
    for d in range(3):
        md.period[d] = float(mesh["period"][d])
    return md


class FiniteElement(System):
    """A DG/FR spatial discretisation on one GPU (fem.hpp:47-436).

    ``scheme``: "fr" (spatial/fr/general.hpp) or "dg" (spatial/dg/general.hpp).
    ``riemann``: "exact", "hllc" or "roe".  ``bcs``: list of (kind, state) with
    kind in {"smart", "inviscid_wall", "supersonic_outlet", "supersonic_inlet"}
    indexed by the mesh's boundary tags (fem.hpp:196-222).
    """

    _RIEMANN = {"exact": C.RIEMANN_EXACT, "hllc": C.RIEMANN_HLLC, "roe": C.RIEMANN_ROE}
    _BC = {"smart": C.BC_SMART, "inviscid_wall": C.BC_INVISCID_WALL,
           "supersonic_outlet": C.BC_SUPERSONIC_OUTLET, "supersonic_inlet": C.BC_SUPERSONIC_INLET,
           "subsonic_inlet": C.BC_SUBSONIC_INLET, "subsonic_outlet": C.BC_SUBSONIC_OUTLET}

    _LIMITER = {"none": C.LIMITER_NONE, "weno": C.LIMITER_WENO_FLAGGED, "weno_all": C.LIMITER_WENO_ALL,
                "weno_eigen": C.LIMITER_WENO_EIGEN, "weno_eigen_all": C.LIMITER_WENO_EIGEN_ALL}

    def __init__(self, mesh: dict, degree: int, scheme: str = "fr", riemann: str = "roe",
                 nodes: str = "legendre", c_next: float = 1.0, bcs=(), precision: str = "fp64",
                 host_only: bool = False, limiter: str = "none", weno_w0: float = 0.001,
                 weno_eps: float = 1e-6):
        lib = C.load()
        self._keep = [np.ascontiguousarray(mesh["coords"], np.float64),
                      np.ascontiguousarray(mesh["cells"], np.int64),
                      np.ascontiguousarray(mesh["bface_nodes"], np.int64),
                      np.ascontiguousarray(mesh["bface_tags"], np.int32)]
        coords, cells, bn, bt = self._keep
        md = C.MeshDesc()
        md.cell_type = int(mesh.get("cell_type", C.HEXA_8))
        md.n_nodes, md.coords = len(coords), C.ptr(coords, C._dp)
        md.n_cells, md.cell_nodes = len(cells), C.ptr(cells, C._i64p)
        md.n_bfaces, md.bface_nodes, md.bface_tags = len(bn), C.ptr(bn, C._i64p), C.ptr(bt, C._i32p)
        for d in range(3):
            md.period[d] = float(mesh["period"][d])
        sd = C.SchemeDesc()
        sd.scheme = C.SCHEME_FR_GENERAL if scheme == "fr" else C.SCHEME_DG_GENERAL
        sd.degree = degree
        sd.nodes = C.NODES_LOBATTO if nodes == "lobatto" else C.NODES_LEGENDRE
        sd.riemann = self._RIEMANN[riemann]
        sd.precision = C.PRECISION_MIXED if precision == "mixed" else C.PRECISION_FP64
        sd.flags = C.FLAG_HOST_ONLY if host_only else 0
        sd.c_next = c_next
        sd.limiter = self._LIMITER[limiter]
        sd.weno_w0, sd.weno_eps = float(weno_w0), float(weno_eps)
        bc_arr = (C.BcDesc * max(1, len(bcs)))()
        for i, (kind, state) in enumerate(bcs):
            bc_arr[i].kind = self._BC[kind]
            for k in range(5):
                bc_arr[i].state[k] = float(state[k]) if state is not None else 0.0
        h = ctypes.c_void_p()
        C.check(lib.mcfd_create(ctypes.byref(md), ctypes.byref(sd), bc_arr, len(bcs), ctypes.byref(h)))
        self._h = h
        n_cells, n_per_cell, size = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
        C.check(lib.mcfd_sizes(h, ctypes.byref(n_cells), ctypes.byref(n_per_cell), ctypes.byref(size)))
        self.n_cells, self.n_per_cell, self.column_size = n_cells.value, n_per_cell.value, size.value
        self.degree = degree
        self.scheme = scheme
        self.cell_type = md.cell_type

    def close(self):
        if getattr(self, "_h", None):
            C.load().mcfd_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- temporal::System
    def set_time(self, t_curr):
        C.check(C.load().mcfd_set_time(self._h, float(t_curr)))

    def set_solution_column(self, column):
        column = np.ascontiguousarray(column, np.float64).reshape(-1)
        assert column.size == self.column_size
        C.check(C.load().mcfd_set_solution_column(self._h, C.ptr(column, C._dp)))

    def restore_solution_column(self, column):
        """Sets the solution as stored, without the limiter pass of
        set_solution_column: the restart path (Part::ReadSolutions +
        ScatterSolutions, demo/euler/sourceless.cpp:167-171)."""
        column = np.ascontiguousarray(column, np.float64).reshape(-1)
        assert column.size == self.column_size
        C.check(C.load().mcfd_restore_solution_column(self._h, C.ptr(column, C._dp)))

    def get_solution_column(self):
        out = np.empty(self.column_size)
        C.check(C.load().mcfd_get_solution_column(self._h, C.ptr(out, C._dp)))
        return out

    def get_residual_column(self):
        out = np.empty(self.column_size)
        C.check(C.load().mcfd_get_residual_column(self._h, C.ptr(out, C._dp)))
        return out

    def boundary_points(self, tag):
        """Points of the faces with boundary tag ``tag``, in the order a boundary function receives them."""
        n = ctypes.c_int64()
        C.check(C.load().mcfd_boundary_points(self._h, int(tag), ctypes.byref(n), None))
        out = np.empty((n.value, 3))
        C.check(C.load().mcfd_boundary_points(self._h, int(tag), ctypes.byref(n), C.ptr(out, C._dp)))
        return out

    def set_boundary_function(self, tag, func):
        """Boundary values f(x, t) of a tag (spatial/fem.hpp:196-222): ``func(xyz[n, 3], t)``
        returns the given states [n, 5] at the tag's points; None restores the constant state."""
        if not hasattr(self, "_bc_callbacks"):
            self._bc_callbacks = {}
        if func is None:
            C.check(C.load().mcfd_set_boundary_function(self._h, int(tag), C.BC_FUNCTION(), None))
            self._bc_callbacks.pop(tag, None)
            return

        def trampoline(n, xyz, t, states, _user):
            x = np.ctypeslib.as_array(xyz, shape=(n, 3))
            out = np.ctypeslib.as_array(states, shape=(n, 5))
            out[...] = func(x.copy(), t)

        cb = C.BC_FUNCTION(trampoline)
        self._bc_callbacks[tag] = cb  # kept alive while the library may call it
        C.check(C.load().mcfd_set_boundary_function(self._h, int(tag), cb, None))

    def set_boundary_values(self, tag, states):
        """Fixed boundary values of a tag at its points (boundary_points(tag) order, [n, 5]):
        a boundary function of the point only, evaluated once."""
        states = np.ascontiguousarray(states, np.float64).reshape(-1, 5)
        assert len(states) == len(self.boundary_points(tag))
        C.check(C.load().mcfd_set_boundary_values(self._h, int(tag), C.ptr(states, C._dp)))
        if hasattr(self, "_bc_callbacks"):
            self._bc_callbacks.pop(tag, None)

    def limiter_reconstruct(self):
        """limiter::Reconstruct (limiter/reconstruct.hpp:22-54): one limiter pass over the current solution."""
        C.check(C.load().mcfd_limiter_reconstruct(self._h))

    def startup_limiter(self, kind):
        """Start-up limiter of a nodal solution: "average" (limiter/average.hpp) or "majority" (majority.hpp)."""
        C.check(C.load().mcfd_startup_limiter(self._h, {"average": C.STARTUP_AVERAGE,
                                                         "majority": C.STARTUP_MAJORITY}[kind]))

    def limiter_flags(self):
        """Troubled-cell flags of the last limiter pass (limiter/reconstruct.hpp:29-34)."""
        flags = np.zeros(self.n_cells, np.uint8)
        n = ctypes.c_int64()
        C.check(C.load().mcfd_limiter_flags(self._h, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                            ctypes.byref(n)))
        return flags

    # ---- fused device paths
    def rk3_update(self, t_curr, dt):
        C.check(C.load().mcfd_rk3_update(self._h, float(t_curr), float(dt)))

    def rk3_stage(self, stage, t_curr, dt):
        C.check(C.load().mcfd_rk3_stage(self._h, int(stage), float(t_curr), float(dt)))

    def synchronize(self):
        C.check(C.load().mcfd_synchronize(self._h))

    def get_time_step(self, cfl, dt_guess=1e300):
        dt = ctypes.c_double()
        C.check(C.load().mcfd_get_time_step(self._h, float(cfl), float(dt_guess), ctypes.byref(dt)))
        return dt.value

    def time_rk3(self, dt, n_steps, per_stage=False):
        """Event-timed fused RK3 steps: (total ms, summed stage-kernel ms[, ms of each stage])."""
        total, kern = ctypes.c_double(), ctypes.c_double()
        stages = np.zeros(3 * int(n_steps))
        C.check(C.load().mcfd_time_rk3(self._h, float(dt), int(n_steps), ctypes.byref(total),
                                       ctypes.byref(kern), C.ptr(stages, C._dp)))
        return (total.value, kern.value, stages) if per_stage else (total.value, kern.value)

    # ---- set-up helpers
    def sample_points(self):
        n = ctypes.c_int32()
        C.check(C.load().mcfd_sample_points(self._h, ctypes.byref(n), None))
        out = np.empty((self.n_cells, n.value, 3))
        C.check(C.load().mcfd_sample_points(self._h, ctypes.byref(n), C.ptr(out, C._dp)))
        return out

    def evaluate(self, local, column=None):
        """(xyz [n_cells, n, 3], values [n_cells, n, 5]) of every cell at the
        local points `local` [n, 3] (polynomial().LocalToGlobalAndValue): the
        current solution, or `column` (which a host-only handle needs)."""
        local = np.ascontiguousarray(local, np.float64).reshape(-1, 3)
        xyz = np.empty((self.n_cells, len(local), 3))
        values = np.empty((self.n_cells, len(local), 5))
        col = None if column is None else np.ascontiguousarray(column, np.float64).reshape(-1)
        C.check(C.load().mcfd_evaluate(self._h, C.ptr(col, C._dp) if col is not None else None, len(local),
                                       C.ptr(local, C._dp), C.ptr(xyz, C._dp), C.ptr(values, C._dp)))
        return xyz, values

    def approximate(self, values):
        """Set the state from values at sample_points(): nodal values (FR) or
        the L2 projection onto each cell's basis (DG, projection.hpp:28-43)."""
        values = np.ascontiguousarray(values, np.float64)
        C.check(C.load().mcfd_approximate(self._h, C.ptr(values, C._dp)))

    def faces(self):
        nf, ni = ctypes.c_int64(), ctypes.c_int64()
        C.check(C.load().mcfd_face_count(self._h, ctypes.byref(nf), ctypes.byref(ni)))
        nf = nf.value
        holder, sharer = np.empty(nf, np.int64), np.empty(nf, np.int64)
        hf, sf = np.empty(nf, np.int32), np.empty(nf, np.int32)
        ijk = np.empty((nf, (self.degree + 1) ** 2, 2), np.int32) if self.scheme == "fr" else None
        C.check(C.load().mcfd_faces(self._h, C.ptr(holder, C._i64p), C.ptr(sharer, C._i64p),
                                    C.ptr(hf, C._i32p), C.ptr(sf, C._i32p),
                                    C.ptr(ijk, C._i32p) if ijk is not None else None))
        out = {"n_interior": ni.value, "holder": holder, "sharer": sharer, "holder_face": hf,
               "sharer_face": sf}
        if ijk is not None:
            out["fp_ijk"] = ijk
        return out


class Batch:
    """Many independent cases advanced by one persistent launch per RK stage.

    ``cases``: list of (mesh, degree, bcs) with bcs as in FiniteElement.  The
    three stages of a step run as a captured hipGraph (``rk3_steps``).
    """

    def __init__(self, cases, riemann="hllc", nodes="legendre", c_next=1.0, precision="fp64"):
        lib = C.load()
        self._keep = []
        n = len(cases)
        descs = (C.MeshDesc * n)()
        degrees = np.array([c[1] for c in cases], dtype=np.int32)
        offsets = np.zeros(n + 1, dtype=np.int32)
        all_bcs = []
        for i, (mesh, degree, bcs) in enumerate(cases):
            descs[i] = _mesh_desc(mesh, self._keep)
            offsets[i + 1] = offsets[i] + len(bcs)
            all_bcs.extend(bcs)
        bc_arr = (C.BcDesc * max(1, len(all_bcs)))()
        for i, (kind, state) in enumerate(all_bcs):
            bc_arr[i].kind = FiniteElement._BC[kind]
            for k in range(5):
                bc_arr[i].state[k] = float(state[k]) if state is not None else 0.0
        sd = C.SchemeDesc()
        sd.scheme = C.SCHEME_FR_GENERAL
        sd.degree = 0
        sd.nodes = C.NODES_LOBATTO if nodes == "lobatto" else C.NODES_LEGENDRE
        sd.riemann = FiniteElement._RIEMANN[riemann]
        sd.precision = C.PRECISION_MIXED if precision == "mixed" else C.PRECISION_FP64
        sd.c_next = c_next
        h = ctypes.c_void_p()
        C.check(lib.mcfd_batch_create(n, descs, C.ptr(degrees, C._i32p), ctypes.byref(sd), bc_arr,
                                      C.ptr(offsets, C._i32p), ctypes.byref(h)))
        self._h = h
        self.n_cases = n
        self.sizes = []
        for i in range(n):
            nc, npc = ctypes.c_int64(), ctypes.c_int32()
            C.check(lib.mcfd_batch_case_sizes(h, i, ctypes.byref(nc), ctypes.byref(npc)))
            self.sizes.append((nc.value, npc.value))

    def close(self):
        if getattr(self, "_h", None):
            C.load().mcfd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def sample_points(self, case):
        nc, npc = self.sizes[case]
        out = np.empty((nc, npc, 3))
        C.check(C.load().mcfd_batch_sample_points(self._h, case, C.ptr(out, C._dp)))
        return out

    def set_solution_column(self, case, column):
        column = np.ascontiguousarray(column, np.float64).reshape(-1)
        nc, npc = self.sizes[case]
        assert column.size == nc * npc * 5
        C.check(C.load().mcfd_batch_set_solution_column(self._h, case, C.ptr(column, C._dp)))

    def get_solution_column(self, case):
        nc, npc = self.sizes[case]
        out = np.empty(nc * npc * 5)
        C.check(C.load().mcfd_batch_get_solution_column(self._h, case, C.ptr(out, C._dp)))
        return out

    def get_time_steps(self, cfl):
        out = np.empty(self.n_cases)
        C.check(C.load().mcfd_batch_get_time_steps(self._h, float(cfl), C.ptr(out, C._dp)))
        return out

    def set_time_steps(self, dt):
        dt = np.ascontiguousarray(dt, np.float64)
        C.check(C.load().mcfd_batch_set_time_steps(self._h, C.ptr(dt, C._dp)))

    def autotune(self, repeats=3):
        """Returns (tiles per degree group, mode): mode 1 = persistent launch."""
        tiles = np.zeros(4, dtype=np.int32)
        mode = ctypes.c_int32()
        C.check(C.load().mcfd_batch_autotune(self._h, int(repeats), C.ptr(tiles, C._i32p), ctypes.byref(mode)))
        return tiles, mode.value

    def set_mode(self, mode, tiles=None):
        """Fix the launch form: mode 1 = persistent launch per stage, 0 = one launch
        per degree group; ``tiles`` = tile variant (0..2) per degree group, ascending degree."""
        t = None if tiles is None else np.ascontiguousarray(tiles, np.int32)
        C.check(C.load().mcfd_batch_set_mode(self._h, int(mode), C.ptr(t, C._i32p) if t is not None else None))

    def rk3_steps(self, n_steps):
        C.check(C.load().mcfd_batch_rk3_steps(self._h, int(n_steps)))

    def synchronize(self):
        C.check(C.load().mcfd_batch_synchronize(self._h))

    def time_rk3(self, n_steps):
        ms = ctypes.c_double()
        C.check(C.load().mcfd_batch_time_rk3(self._h, int(n_steps), ctypes.byref(ms)))
        return ms.value
