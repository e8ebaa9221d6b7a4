"""ctypes loader of the oracle (oracle/_build/liboracle.so): test infrastructure.

Only tests, the smoke check and bench.py's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_lib = None


def _ptr(a, t):
    return a.ctypes.data_as(t)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_fr_create.argtypes = [ctypes.c_int64, _dp, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,
                                     _i32p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, _i32p, _dp, ctypes.POINTER(ctypes.c_void_p)]
    lib.oracle_fr_n_faces.restype = ctypes.c_int64
    lib.oracle_fr_n_faces.argtypes = [ctypes.c_void_p, _i64p]
    lib.oracle_fr_faces.argtypes = [ctypes.c_void_p, _i64p, _i64p, _i32p, _i32p, _i32p, _dp]
    lib.oracle_fr_node_coords.argtypes = [ctypes.c_void_p, _dp]
    lib.oracle_fr_node_weights.argtypes = [ctypes.c_void_p, _dp]
    lib.oracle_fr_cell_lengths.argtypes = [ctypes.c_void_p, _dp]
    lib.oracle_fr_residual.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.oracle_fr_rk3.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int]
    lib.oracle_fr_time_step.restype = ctypes.c_double
    lib.oracle_fr_time_step.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double]
    lib.oracle_fr_time_residuals.restype = ctypes.c_double
    lib.oracle_fr_time_residuals.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    lib.oracle_fr_destroy.argtypes = [ctypes.c_void_p]
    lib.oracle_quadrature.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.oracle_vincent.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
    lib.oracle_lagrange.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
    for name in ("oracle_conservative_to_primitive", "oracle_primitive_to_conservative",
                 "oracle_flux_matrix"):
        getattr(lib, name).argtypes = [_dp, _dp]
    lib.oracle_exact_primitive.argtypes = [_dp, _dp, _dp]
    lib.oracle_riemann_flux.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.oracle_wall_flux.argtypes = [_dp, _dp, _dp]
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(load().oracle_last_error().decode())


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def quadrature(lobatto, q):
    x, w = np.zeros(q), np.zeros(q)
    _check(load().oracle_quadrature(int(lobatto), q, _ptr(x, _dp), _ptr(w, _dp)))
    return x, w


def vincent(degree, c_next, s):
    out = np.zeros(4)
    load().oracle_vincent(degree, c_next, s, _ptr(out, _dp))
    return out


def lagrange(x, s):
    x = f64(x)
    v, d = np.zeros(len(x)), np.zeros(len(x))
    load().oracle_lagrange(len(x), _ptr(x, _dp), s, _ptr(v, _dp), _ptr(d, _dp))
    return v, d


def conservative_to_primitive(cv):
    cv, pv = f64(cv), np.zeros(5)
    load().oracle_conservative_to_primitive(_ptr(cv, _dp), _ptr(pv, _dp))
    return pv


def primitive_to_conservative(pv):
    pv, cv = f64(pv), np.zeros(5)
    load().oracle_primitive_to_conservative(_ptr(pv, _dp), _ptr(cv, _dp))
    return cv


def flux_matrix(cv):
    cv, f = f64(cv), np.zeros(15)
    load().oracle_flux_matrix(_ptr(cv, _dp), _ptr(f, _dp))
    return f.reshape(5, 3)


def exact_primitive(pl, pr):
    pl, pr, out = f64(pl), f64(pr), np.zeros(5)
    load().oracle_exact_primitive(_ptr(pl, _dp), _ptr(pr, _dp), _ptr(out, _dp))
    return out


def riemann_flux(kind, frame, cl, cr):
    frame, cl, cr, out = f64(frame).reshape(9), f64(cl), f64(cr), np.zeros(5)
    load().oracle_riemann_flux(kind, _ptr(frame, _dp), _ptr(cl, _dp), _ptr(cr, _dp), _ptr(out, _dp))
    return out


def face_fluxes(riemann, kind, frames, left, right=None, star_u=None):
    """Fluxes at n points, each with its own solver; star_u (in/out) is each
    point's exact-solver state (exact.hpp:29)."""
    lib = load()
    lib.oracle_face_fluxes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp]
    frames, left = f64(frames).reshape(-1, 9), f64(left).reshape(-1, 5)
    n = len(left)
    right = f64(right).reshape(-1, 5) if right is not None else None
    out = np.zeros((n, 5))
    lib.oracle_face_fluxes(riemann, kind, n, _ptr(frames, _dp), _ptr(left, _dp),
                           _ptr(right, _dp) if right is not None else None,
                           _ptr(star_u, _dp) if star_u is not None else None, _ptr(out, _dp))
    return out


def wall_flux(frame, c):
    frame, c, out = f64(frame).reshape(9), f64(c), np.zeros(5)
    load().oracle_wall_flux(_ptr(frame, _dp), _ptr(c, _dp), _ptr(out, _dp))
    return out


def eigen_matrices(cv, frame):
    """riemann/euler/eigen.hpp: (L, R) of state cv in the frame rows (n, t, b)."""
    lib = load()
    lib.oracle_eigen_matrices.argtypes = [_dp, _dp, _dp, _dp]
    cv, frame = f64(cv), f64(frame).reshape(9)
    l, r = np.zeros((5, 5)), np.zeros((5, 5))
    lib.oracle_eigen_matrices(_ptr(cv, _dp), _ptr(frame, _dp), _ptr(l, _dp), _ptr(r, _dp))
    return l, r


class _BoundaryValues:
    """Boundary values that vary in space and time (spatial/fem.hpp:196-222):
    the caller's functions f(xyz, t) evaluated at the restatement's boundary
    points, the other tags' constant states elsewhere."""

    _prefix = ""

    def boundary_points(self):
        """(xyz [n, 3], tag [n]) of every boundary face point, in face order."""
        lib = load()
        fn = getattr(lib, f"oracle_{self._prefix}_boundary_points")
        fn.restype = ctypes.c_int64
        fn.argtypes = [ctypes.c_void_p, _dp, _i32p]
        n = fn(self.h, None, None)
        xyz, tags = np.zeros((n, 3)), np.zeros(n, np.int32)
        fn(self.h, _ptr(xyz, _dp), _ptr(tags, _i32p))
        return xyz, tags

    def set_boundary_states(self, states):
        """Given state at every boundary point ([n, 5]), or None for the tags' constants."""
        lib = load()
        fn = getattr(lib, f"oracle_{self._prefix}_set_boundary_states")
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, _dp]
        if states is None:
            fn(self.h, 0, None)
            return
        states = f64(states).reshape(-1, 5)
        fn(self.h, len(states), _ptr(states, _dp))

    def boundary_states_at(self, t, functions, constants):
        """Given states at time t: functions {tag: f(xyz, t) -> [n, 5]}, constants [tag][5]."""
        xyz, tags = self.boundary_points()
        out = np.asarray(constants, dtype=np.float64)[tags].copy()
        for tag, f in functions.items():
            m = tags == tag
            if m.any():
                out[m] = f(xyz[m], t)
        return out


def rk3_timed(o, u, t, dt, functions, constants, limit=None):
    """One SSP-RK3 step of temporal/rk.hpp:31-49 (operation order of
    ode.hpp:121-128) with the boundary functions evaluated at the stage times
    t, t + dt and t + dt / 2; ``limit`` is the SetSolutionColumn side effect."""
    def next_solution(state, ts):
        o.set_boundary_states(o.boundary_states_at(ts, functions, constants))
        r = o.residual(state)
        r *= dt
        return state + r

    def set_solution(v):
        return limit(v) if limit is not None else v

    u = f64(u)
    u1 = set_solution(next_solution(u, t))
    u_next = u.copy()
    u_next *= 3
    u_next += next_solution(u1, t + dt)
    u_next /= 4
    u2 = set_solution(u_next)
    v = next_solution(u2, t + dt / 2)
    v *= 2
    v += u
    v /= 3
    return set_solution(v)


class FrOracle(_BoundaryValues):
    """Nodal FR on hexahedra (spatial/fr/general.hpp restated)."""

    _prefix = "fr"

    def __init__(self, mesh: dict, q: int, riemann: int, lobatto=False, c_next=1.0, bcs=()):
        lib = load()
        self.mesh = mesh
        self.q = q
        self.n = q ** 3
        self.n_cells = len(mesh["cells"])
        self._keep = [f64(mesh["coords"]), np.ascontiguousarray(mesh["cells"], np.int64),
                      np.ascontiguousarray(mesh["bface_nodes"], np.int64),
                      np.ascontiguousarray(mesh["bface_tags"], np.int32), f64(mesh["period"])]
        kinds = np.array([b[0] for b in bcs] or [0], dtype=np.int32)
        states = f64(np.array([b[1] for b in bcs] or [[0.0] * 5]).reshape(-1))
        self._keep += [kinds, states]
        h = ctypes.c_void_p()
        c, cells, bn, bt, per = self._keep[:5]
        _check(lib.oracle_fr_create(len(c), _ptr(c, _dp), len(cells), _ptr(cells, _i64p), len(bn),
                                    _ptr(bn, _i64p), _ptr(bt, _i32p), _ptr(per, _dp), q, int(lobatto),
                                    float(c_next), int(riemann), len(bcs), _ptr(kinds, _i32p),
                                    _ptr(states, _dp), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_fr_destroy(self.h)
            self.h = None

    @property
    def size(self):
        return self.n_cells * self.n * 5

    def node_coords(self):
        out = np.zeros((self.n_cells, self.n, 3))
        load().oracle_fr_node_coords(self.h, _ptr(out, _dp))
        return out

    def node_weights(self):
        out = np.zeros((self.n_cells, self.n))
        load().oracle_fr_node_weights(self.h, _ptr(out, _dp))
        return out

    def cell_lengths(self):
        out = np.zeros(self.n_cells)
        load().oracle_fr_cell_lengths(self.h, _ptr(out, _dp))
        return out

    def faces(self):
        n_int = ctypes.c_int64()
        nf = load().oracle_fr_n_faces(self.h, ctypes.byref(n_int))
        nfp = self.q * self.q
        holder, sharer = np.zeros(nf, np.int64), np.zeros(nf, np.int64)
        hf, sf = np.zeros(nf, np.int32), np.zeros(nf, np.int32)
        ijk = np.zeros((nf, nfp, 2), np.int32)
        frames = np.zeros((nf, nfp, 3, 3))
        load().oracle_fr_faces(self.h, _ptr(holder, _i64p), _ptr(sharer, _i64p), _ptr(hf, _i32p),
                               _ptr(sf, _i32p), _ptr(ijk, _i32p), _ptr(frames, _dp))
        return {"n_interior": n_int.value, "holder": holder, "sharer": sharer, "holder_face": hf,
                "sharer_face": sf, "fp_ijk": ijk, "frames": frames}

    def residual(self, u):
        u = f64(u)
        r = np.zeros_like(u)
        _check(load().oracle_fr_residual(self.h, _ptr(u, _dp), _ptr(r, _dp)))
        return r

    def startup_limiter(self, u, kind):
        """limiter/average.hpp ("average") or limiter/majority.hpp ("majority") on a nodal column."""
        lib = load()
        lib.oracle_fr_startup_limiter.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        u = f64(u).copy()
        _check(lib.oracle_fr_startup_limiter(self.h, _ptr(u, _dp), {"average": 0, "majority": 1}[kind]))
        return u

    def set_mutation(self, mutation: int):
        """Plants a known mistake in the restatement (0 = none); see FrScheme::mutation."""
        lib = load()
        lib.oracle_fr_set_mutation.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.oracle_fr_set_mutation(self.h, int(mutation))

    def rk3(self, u, dt, n_steps=1):
        u = f64(u).copy()
        _check(load().oracle_fr_rk3(self.h, _ptr(u, _dp), float(dt), int(n_steps)))
        return u

    def rk3_threaded(self, u, dt, n_steps=1, threads=4):
        """rk3 with the residual spread over host threads (bitwise the same states)."""
        lib = load()
        lib.oracle_fr_rk3_threaded.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        u = f64(u).copy()
        _check(lib.oracle_fr_rk3_threaded(self.h, _ptr(u, _dp), float(dt), int(n_steps), int(threads)))
        return u

    def time_step(self, u, cfl, dt_guess=1e300):
        u = f64(u)
        return load().oracle_fr_time_step(self.h, _ptr(u, _dp), cfl, dt_guess)

    def time_residuals(self, u, n):
        u = f64(u)
        r = np.zeros_like(u)
        return load().oracle_fr_time_residuals(self.h, _ptr(u, _dp), _ptr(r, _dp), n)


class DgOracle(_BoundaryValues):
    """Modal DG on hexahedra or tetrahedra (spatial/dg/general.hpp restated)."""

    _prefix = "dg"

    def __init__(self, mesh: dict, degree: int, riemann: int, bcs=()):
        lib = load()
        if not getattr(lib, "_dg_bound", False):
            lib.oracle_dg_create.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_int64, _i64p,
                                             ctypes.c_int64, _i64p, _i32p, _dp, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _i32p, _dp, ctypes.POINTER(ctypes.c_void_p)]
            lib.oracle_dg_destroy.argtypes = [ctypes.c_void_p]
            lib.oracle_dg_sizes.argtypes = [ctypes.c_void_p, _i64p]
            lib.oracle_dg_cells.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
            lib.oracle_dg_faces.argtypes = [ctypes.c_void_p, _i64p, _i64p, _i32p, _i32p, _i32p, _dp, _dp, _dp]
            lib.oracle_dg_residual.argtypes = [ctypes.c_void_p, _dp, _dp]
            lib.oracle_dg_rk3.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int]
            lib.oracle_dg_time_step.restype = ctypes.c_double
            lib.oracle_dg_time_step.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double]
            lib.oracle_dg_approximate.argtypes = [ctypes.c_void_p, _dp, _dp]
            lib.oracle_dg_value.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, _dp, _dp]
            lib._dg_bound = True
        self.mesh = mesh
        cells = np.ascontiguousarray(mesh["cells"], np.int64)
        self._keep = [f64(mesh["coords"]), cells, np.ascontiguousarray(mesh["bface_nodes"], np.int64),
                      np.ascontiguousarray(mesh["bface_tags"], np.int32), f64(mesh["period"])]
        kinds = np.array([b[0] for b in bcs] or [0], dtype=np.int32)
        states = f64(np.array([b[1] for b in bcs] or [[0.0] * 5]).reshape(-1))
        self._keep += [kinds, states]
        c, cl, bn, bt, per = self._keep[:5]
        h = ctypes.c_void_p()
        _check(lib.oracle_dg_create(int(mesh.get("cell_type", 0)), len(c), _ptr(c, _dp), len(cl), _ptr(cl, _i64p),
                                    len(bn), _ptr(bn, _i64p), _ptr(bt, _i32p), _ptr(per, _dp), degree,
                                    int(riemann), len(bcs), _ptr(kinds, _i32p), _ptr(states, _dp),
                                    ctypes.byref(h)))
        self.h = h
        sizes = np.zeros(5, np.int64)
        lib.oracle_dg_sizes(h, _ptr(sizes, _i64p))
        self.n_cells, self.n, self.nq, self.n_faces, self.n_interior = (int(v) for v in sizes)
        self.degree = degree

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_dg_destroy(self.h)
            self.h = None

    @property
    def size(self):
        return self.n_cells * self.n * 5

    def cells(self):
        xyz = np.zeros((self.n_cells, self.nq, 3))
        w = np.zeros((self.n_cells, self.nq))
        centers = np.zeros((self.n_cells, 3))
        s = np.zeros((self.n_cells, self.n, self.n))
        lengths = np.zeros(self.n_cells)
        load().oracle_dg_cells(self.h, _ptr(xyz, _dp), _ptr(w, _dp), _ptr(centers, _dp), _ptr(s, _dp),
                               _ptr(lengths, _dp))
        return {"xyz": xyz, "w": w, "centers": centers, "s": s, "lengths": lengths}

    def faces(self):
        nf = self.n_faces
        nfp = 6 if self.mesh.get("cell_type", 0) == 1 else (self.degree + 1) ** 2
        holder, sharer = np.zeros(nf, np.int64), np.zeros(nf, np.int64)
        hf, sf, tag = np.zeros(nf, np.int32), np.zeros(nf, np.int32), np.zeros(nf, np.int32)
        xyz, w, frames = np.zeros((nf, nfp, 3)), np.zeros((nf, nfp)), np.zeros((nf, nfp, 3, 3))
        load().oracle_dg_faces(self.h, _ptr(holder, _i64p), _ptr(sharer, _i64p), _ptr(hf, _i32p),
                               _ptr(sf, _i32p), _ptr(tag, _i32p), _ptr(xyz, _dp), _ptr(w, _dp), _ptr(frames, _dp))
        return {"n_interior": self.n_interior, "holder": holder, "sharer": sharer, "holder_face": hf,
                "sharer_face": sf, "tag": tag, "xyz": xyz, "w": w, "frames": frames}

    def residual(self, u):
        u = f64(u)
        r = np.zeros_like(u)
        _check(load().oracle_dg_residual(self.h, _ptr(u, _dp), _ptr(r, _dp)))
        return r

    def rk3(self, u, dt, n_steps=1):
        u = f64(u).copy()
        _check(load().oracle_dg_rk3(self.h, _ptr(u, _dp), float(dt), int(n_steps)))
        return u

    def time_step(self, u, cfl, dt_guess=1e300):
        u = f64(u)
        return load().oracle_dg_time_step(self.h, _ptr(u, _dp), cfl, dt_guess)

    def approximate(self, values):
        """values[cell][q][5] at the cell quadrature points -> coefficients."""
        values = f64(values)
        u = np.zeros(self.size)
        load().oracle_dg_approximate(self.h, _ptr(values, _dp), _ptr(u, _dp))
        return u

    def value(self, u, cell, x):
        u, x, v = f64(u), f64(x), np.zeros(5)
        load().oracle_dg_value(self.h, _ptr(u, _dp), int(cell), _ptr(x, _dp), _ptr(v, _dp))
        return v

    # ---- WENO limiter (limiter/weno.hpp restated)
    def set_limiter(self, mode: int, w0: float = 0.001, eps: float = 1e-6):
        lib = load()
        lib.oracle_dg_set_limiter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        lib.oracle_dg_set_limiter(self.h, int(mode), float(w0), float(eps))

    def limit(self, u):
        """Applies the limiter (as SetSolutionColumn does); returns (new u, flags)."""
        lib = load()
        lib.oracle_dg_limit.restype = ctypes.c_int64
        lib.oracle_dg_limit.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_uint8)]
        u = f64(u).copy()
        flags = np.zeros(self.n_cells, np.uint8)
        lib.oracle_dg_limit(self.h, _ptr(u, _dp), flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return u, flags

    def last_flags(self):
        """Flags of the oracle's last limiter pass (the last RK stage after rk3)."""
        lib = load()
        lib.oracle_dg_last_flags.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8)]
        flags = np.zeros(self.n_cells, np.uint8)
        lib.oracle_dg_last_flags(self.h, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return flags

    def smoothness(self, cell, coeff):
        lib = load()
        lib.oracle_dg_smoothness.argtypes = [ctypes.c_void_p, ctypes.c_int64, _dp, _dp]
        coeff, beta = f64(coeff), np.zeros(5)
        lib.oracle_dg_smoothness(self.h, int(cell), _ptr(coeff, _dp), _ptr(beta, _dp))
        return beta
