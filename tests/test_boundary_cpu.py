"""CPU checks of the C ABI and of the host-side logic (no GPU needed).

* libminicfd exports every function include/minicfd.h declares, and the
  ctypes table binds each of them;
* a host-only handle (MCFD_FLAG_HOST_ONLY) builds the connectivity without
  touching a device: holder/sharer ids, local face ids and the face-point to
  solution-node maps must equal the oracle's bit for bit (mesh/part.hpp:932-975,
  1456-1549; coordinate/cell.hpp:235-258), including cells whose vertex lists
  start at any of the 24 rotations of the reference hexahedron;
* the error contract: bad arguments give codes and messages, device calls on a
  host-only handle are refused;
* temporal.RungeKutta reproduces test/temporal/rk.cpp:53-70 bit for bit.
"""
from __future__ import annotations

import ctypes
import itertools
import os
import re

import numpy as np
import pytest

import oracle_lib as ol
from pvcriemannsolvers_amd import _capi as C
from pvcriemannsolvers_amd import meshgen, spatial, temporal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "minicfd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(mcfd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(C.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(C.SIGNATURES) == names


def test_last_error_after_bad_arguments():
    lib = C.load()
    h = ctypes.c_void_p()
    assert lib.mcfd_create(None, None, None, 0, ctypes.byref(h)) == C.MCFD_EINVAL
    assert lib.mcfd_last_error().decode() == "null argument"
    mesh = meshgen.box_hex(2, 2, 2)
    with pytest.raises(C.MinicfdError) as e:
        spatial.FiniteElement(mesh, degree=9, host_only=True)
    assert e.value.code == C.MCFD_EINVAL
    with pytest.raises(C.MinicfdError) as e:  # a boundary tag without a BC entry
        spatial.FiniteElement(mesh, degree=2, host_only=True, bcs=[("smart", [1, 0, 0, 0, 2.5])])
    assert e.value.code == C.MCFD_EINVAL


def test_host_only_handle_refuses_device_calls():
    fe = spatial.FiniteElement(meshgen.box_hex(2, 2, 2, periodic=(True,) * 3), degree=1, host_only=True)
    assert fe.column_size == 8 * 8 * 5
    with pytest.raises(C.MinicfdError) as e:
        fe.get_residual_column()
    assert e.value.code == C.MCFD_ESTATE
    fe.close()


# ---------------------------------------------------------------- connectivity
_CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]) * 2 - 1


def cube_rotations():
    """The 24 proper rotations of the cube as vertex relabelings of a HEXA_8."""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            r = np.zeros((3, 3), int)
            for i, (p, s) in enumerate(zip(perm, signs)):
                r[i, p] = s
            if round(np.linalg.det(r)) != 1:
                continue
            moved = _CORNERS @ r.T
            out.append([int(np.flatnonzero((_CORNERS == m).all(1))[0]) for m in moved])
    assert len(out) == 24
    return out


def rotate_cells(mesh, seed):
    rots = cube_rotations()
    rng = np.random.default_rng(seed)
    cells = mesh["cells"].copy()
    pick = rng.integers(0, 24, len(cells))
    for c in range(len(cells)):
        cells[c] = cells[c][rots[pick[c]]]
    return dict(mesh, cells=cells)


WALLS = [("smart", [1.0, 0.0, 0.0, 0.0, 2.5]), ("smart", [0.125, 0.0, 0.0, 0.0, 0.25])] + \
        [("inviscid_wall", None)] * 4
ORACLE_WALLS = [(0, [1.0, 0.0, 0.0, 0.0, 2.5]), (0, [0.125, 0.0, 0.0, 0.0, 0.25])] + [(1, [0.0] * 5)] * 4


def sheared(mesh):
    a = np.array([[1.0, 0.2, 0.1], [0.0, 0.9, 0.3], [0.05, 0.0, 1.1]])
    return dict(mesh, coords=np.ascontiguousarray(mesh["coords"] @ a.T + [0.3, -0.1, 0.2]))


MESHES = {
    "shuffled_walls": (lambda: meshgen.box_hex(6, 3, 2, hi=(1.0, 0.5, 0.3), shuffle_seed=5), True),
    "periodic": (lambda: meshgen.box_hex(4, 3, 5, periodic=(True,) * 3, shuffle_seed=1), False),
    "mixed_periodic": (lambda: meshgen.box_hex(3, 4, 2, periodic=(True, False, True), shuffle_seed=3), True),
    "rotated_cells_walls": (lambda: rotate_cells(meshgen.box_hex(4, 3, 3, shuffle_seed=9), 11), True),
    "rotated_cells_periodic": (lambda: rotate_cells(meshgen.box_hex(3, 3, 3, periodic=(True,) * 3), 12), False),
    "sheared_rotated": (lambda: sheared(rotate_cells(meshgen.box_hex(3, 2, 4, shuffle_seed=2), 13)), True),
    "jittered_rotated": (lambda: rotate_cells(meshgen.box_hex(3, 3, 3, shuffle_seed=6, jitter=0.2, seed=10), 14),
                         True),
    "jittered_periodic": (lambda: meshgen.box_hex(3, 3, 3, periodic=(True,) * 3, jitter=0.2, seed=11), False),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_connectivity_matches_oracle_bit_exact(name, degree):
    make, walls = MESHES[name]
    mesh = make()
    bcs = WALLS if walls else ()
    fe = spatial.FiniteElement(mesh, degree=degree, riemann="hllc", bcs=bcs, host_only=True)
    o = ol.FrOracle(mesh, q=degree + 1, riemann=1, bcs=ORACLE_WALLS if walls else ())
    got, want = fe.faces(), o.faces()
    assert got["n_interior"] == want["n_interior"]
    for key in ("holder", "sharer", "holder_face", "sharer_face", "fp_ijk"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    # every interior face appears once, boundary faces last
    n_int = want["n_interior"]
    sides = list(zip(want["holder"][:n_int], want["holder_face"][:n_int])) + \
        list(zip(want["sharer"][:n_int], want["sharer_face"][:n_int]))
    assert len(set(sides)) == 2 * n_int  # a (cell, local face) is used once
    assert np.all(want["sharer"][n_int:] == -1) and np.all(want["sharer"][:n_int] >= 0)
    # 6 faces per cell
    counts = np.bincount(np.concatenate([want["holder"], want["sharer"][:n_int]]), minlength=fe.n_cells)
    assert np.all(counts == 6)
    # solution points coincide with the oracle's
    xyz = fe.sample_points()
    np.testing.assert_allclose(xyz.reshape(-1, 3), o.node_coords().reshape(-1, 3), rtol=0, atol=1e-14)
    fe.close()


def test_face_points_coincide_geometrically():
    """fem.hpp:43-45: the holder's and sharer's face points are the same global points."""
    mesh = rotate_cells(meshgen.box_hex(3, 3, 2, shuffle_seed=4), 21)
    o = ol.FrOracle(mesh, q=3, riemann=1, bcs=ORACLE_WALLS)
    f = o.faces()
    q = 3
    x, _ = ol.quadrature(False, q)
    xyz = o.node_coords()
    lo = ol.lagrange(x, -1.0)[0]
    hi = ol.lagrange(x, 1.0)[0]
    # face-normal direction and side of each local face (HEXA_8 faces 0..5)
    axis, side = [2, 1, 0, 1, 0, 2], [0, 0, 1, 1, 0, 1]

    def trace(cell, face, node):
        i, j, k = node // (q * q), (node // q) % q, node % q
        idx = [i, j, k]
        ends = hi if side[face] else lo
        pts = []
        for m in range(q):
            idx[axis[face]] = m
            pts.append(xyz[cell, idx[0] * q * q + idx[1] * q + idx[2]])
        return ends @ np.array(pts)

    for fid in range(f["n_interior"]):
        for p in range(q * q):
            a = trace(f["holder"][fid], f["holder_face"][fid], f["fp_ijk"][fid, p, 0])
            b = trace(f["sharer"][fid], f["sharer_face"][fid], f["fp_ijk"][fid, p, 1])
            np.testing.assert_allclose(a, b, atol=1e-13)


# ---------------------------------------------------------------- time stepping (test/temporal/rk.cpp)
@pytest.mark.parametrize("orders", [1, 2, 3])
def test_runge_kutta_bitwise(orders):
    rng = np.random.default_rng(31415926)
    n = 10
    a = rng.uniform(-1, 1, (n, n))
    u_old = rng.uniform(-1, 1, n)
    system = temporal.Constant(a)
    system.set_solution_column(u_old)
    t, dt = rng.random(), rng.random()
    temporal.RungeKutta(orders).update(system, t, dt)
    u1 = u_old + (a @ u_old) * dt
    if orders == 1:
        want = u1
    elif orders == 2:
        want = ((u1 + (a @ u1) * dt) + u_old) * 0.5
    else:
        u2 = ((u1 + (a @ u1) * dt) + u_old * 3) / 4
        want = ((u2 + (a @ u2) * dt) * 2 + u_old) / 3
    assert np.array_equal(system.get_solution_column(), want)


def test_meshgen_box_counts():
    m = meshgen.box_hex(40, 4, 4, hi=(1.0, 0.1, 0.1), shuffle_seed=0)
    assert m["cells"].shape == (640, 8)
    assert len(m["bface_nodes"]) == 2 * (16 + 160 + 160)
    o = ol.FrOracle(m, q=3, riemann=0, bcs=ORACLE_WALLS)
    f = o.faces()
    assert f["n_interior"] == 1584 and len(f["holder"]) - f["n_interior"] == 672


# ---------------------------------------------------------------- modal DG handles
DG_MESHES = {
    "tets_jittered": (lambda: meshgen.box_tet(4, 3, 2, hi=(1.0, 0.6, 0.4), jitter=0.15, seed=7), True),
    "hex_shuffled_walls": (lambda: meshgen.box_hex(5, 3, 2, hi=(1.0, 0.5, 0.3), shuffle_seed=6), True),
    "hex_rotated_periodic": (lambda: rotate_cells(meshgen.box_hex(3, 3, 2, periodic=(True,) * 3), 14), False),
}


@pytest.mark.parametrize("name", sorted(DG_MESHES))
@pytest.mark.parametrize("degree", [1, 2])
def test_dg_connectivity_and_points_match_oracle(name, degree):
    make, walls = DG_MESHES[name]
    mesh = make()
    bcs = WALLS if walls else ()
    fe = spatial.FiniteElement(mesh, degree=degree, scheme="dg", riemann="hllc", bcs=bcs, host_only=True)
    o = ol.DgOracle(mesh, degree=degree, riemann=1, bcs=ORACLE_WALLS if walls else ())
    got, want = fe.faces(), o.faces()
    assert got["n_interior"] == want["n_interior"]
    for key in ("holder", "sharer", "holder_face", "sharer_face"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    assert fe.n_per_cell == o.n and fe.column_size == o.size
    c = o.cells()
    np.testing.assert_allclose(fe.sample_points(), c["xyz"], rtol=0, atol=1e-14)
    fe.close()


def test_unsupported_configurations_are_rejected():
    """Configurations outside the path fail at creation with a status code, not later on the device."""
    walls = [("inviscid_wall", None)] * 6
    trilinear = meshgen.box_hex(2, 2, 2)
    trilinear["coords"] = trilinear["coords"].copy()
    trilinear["coords"][13] += (0.05, -0.03, 0.02)  # the interior vertex of the 2 x 2 x 2 box
    cases = [
        # the WENO limiter needs the modal DG scheme
        (dict(mesh=meshgen.box_hex(2, 2, 2), degree=2, scheme="fr", bcs=walls, limiter="weno"),
         C.MCFD_ENOTSUP, "DG_GENERAL"),
        # and no neighbour polynomial is defined across periodic faces
        (dict(mesh=meshgen.box_hex(3, 3, 3, periodic=(True,) * 3), degree=2, scheme="dg", limiter="weno"),
         C.MCFD_EINVAL, "periodic"),
        # modal DG carries p <= 2
        (dict(mesh=meshgen.box_tet(2, 2, 2), degree=3, scheme="dg", bcs=walls), C.MCFD_EINVAL, "degree"),
        # FR carries Q <= 6 points per direction
        (dict(mesh=meshgen.box_hex(2, 2, 2), degree=6, scheme="fr", bcs=walls), C.MCFD_EINVAL, "degree"),
        # trilinear (non-parallelepiped) hexahedra run with FR only
        (dict(mesh=trilinear, degree=2, scheme="dg", bcs=walls), C.MCFD_ENOTSUP, "affine"),
        (dict(mesh=trilinear, degree=2, bcs=walls, precision="mixed"), C.MCFD_ENOTSUP, "affine"),
    ]
    for kw, code, words in cases:
        mesh = kw.pop("mesh")
        with pytest.raises(C.MinicfdError) as e:
            spatial.FiniteElement(mesh, host_only=True, **kw)
        assert e.value.code == code, kw
        assert words in C.load().mcfd_last_error().decode(), kw
