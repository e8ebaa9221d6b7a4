"""FR on trilinear hexahedra (polynomial/hexahedron.hpp:172-239 per-node J*,
fr/general.hpp:103-205 flux-point scales) against the oracle.

The meshes are boxes whose interior vertices are jittered by +-20 % of the
spacing, so every cell is a general trilinear hexahedron with non-planar
faces: J* and det(J) vary from node to node, the flux-point scales and the
Riemann frames from face point to face point.  The oracle's FR on such cells
is pinned independently by the nodal-DG weak form (tests/test_oracle_weak_form.py,
the "jittered" meshes).  Tolerances as in tests/test_fr_gpu.py: 1e-12
relative (Roe, HLLC), 1e-7 for the exact solver's stopping rule.
"""
import numpy as np
import pytest

import oracle_lib as ol
from pvcriemannsolvers_amd import meshgen, spatial
from test_fr_gpu import TOL, conservative, density_wave, make_pair, rel


def periodic_mesh(seed=7, shuffle=None):
    return meshgen.box_hex(4, 3, 3, periodic=(True,) * 3, jitter=0.2, seed=seed, shuffle_seed=shuffle)


def bounded_mesh():
    return meshgen.box_hex(5, 3, 2, hi=(1.0, 0.3, 0.2), jitter=0.2, seed=8, shuffle_seed=9)


def test_curved_node_and_boundary_points_match_oracle():
    """Host-only: the device layout's solution points and boundary flux points
    are the oracle's trilinear-map images of the same local points."""
    mesh = bounded_mesh()
    bcs = [("smart", conservative(1.0, (0.1, 0, 0), 1.0))] * 2 + [("inviscid_wall", None)] * 4
    g = spatial.FiniteElement(mesh, degree=3, riemann="hllc", bcs=bcs, host_only=True)
    o = ol.FrOracle(mesh, q=4, riemann=1, bcs=[(0 if k == "smart" else 1, s if s is not None else [0] * 5)
                                               for k, s in bcs])
    assert np.abs(g.sample_points() - o.node_coords()).max() < 1e-15
    xyz, tags = o.boundary_points()

    def ordered(p):  # the two layouts order faces and points differently
        p = np.asarray(p).reshape(-1, 3)
        return p[np.lexsort(np.round(p, 12).T[::-1])]
    for tag in range(6):
        got, want = ordered(g.boundary_points(tag)), ordered(xyz[tags == tag])
        assert got.shape == want.shape and np.abs(got - want).max() < 1e-15, tag
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree", [1, 2, 3, 4])
@pytest.mark.parametrize("riemann", ["roe", "hllc", "exact"])
def test_curved_periodic_residual(degree, riemann):
    mesh = periodic_mesh(shuffle=degree)
    o, g = make_pair(mesh, degree, riemann)
    u = density_wave(o.node_coords())
    g.set_solution_column(u.reshape(-1))
    assert rel(g.get_residual_column(), o.residual(u).reshape(-1)) < TOL[riemann]
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("riemann", ["hllc", "roe"])
def test_curved_bounded_residual(degree, riemann):
    """Smart ends, an inviscid wall pair, a supersonic outlet and inlet pair."""
    left, right = conservative(1.0, (0.1, 0, 0), 1.0), conservative(0.125, (0, 0, 0), 0.1)
    bcs = [("smart", left), ("smart", right), ("inviscid_wall", None), ("inviscid_wall", None),
           ("supersonic_outlet", None), ("supersonic_inlet", left)]
    o, g = make_pair(bounded_mesh(), degree, riemann, bcs)
    # the jump smoothed over about a cell: no jittered face sits on x = 0.5,
    # and a jump inside a cell extrapolates to negative face densities
    x = o.node_coords()
    w = 0.5 * (1 + np.tanh((x[..., 0] - 0.5) / 0.15))
    u = ((1 - w)[..., None] * left + w[..., None] * right) * (1 + 1e-3 * np.random.default_rng(2).standard_normal(w.shape))[..., None]
    g.set_solution_column(u.reshape(-1))
    assert rel(g.get_residual_column(), o.residual(u).reshape(-1)) < TOL[riemann]
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree", [2, 3])
def test_curved_lumped_lobatto_residual(degree):
    """fr::Lobatto (c_next = p / (2p + 1) on Lobatto nodes, i.e. DGSEM) on trilinear cells."""
    o, g = make_pair(periodic_mesh(seed=9), degree, "roe", nodes="lobatto", c_next=degree / (2 * degree + 1))
    u = density_wave(o.node_coords())
    g.set_solution_column(u.reshape(-1))
    assert rel(g.get_residual_column(), o.residual(u).reshape(-1)) < 1e-12
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree,riemann", [(3, "exact"), (5, "hllc")])
def test_curved_partial_tile_largest_degree(degree, riemann):
    """Trilinear cells: 45 cells leave the last 2-cell workgroup (Q = 4) half
    empty, Q = 6 runs one cell per workgroup; residual and RK3 steps against the
    oracle.  (At Q = 6 the exact solver's stopping rule, f >= -1e-8, moves the
    residual by about 1e-7 of its norm, so that degree runs with HLLC.)"""
    mesh = meshgen.box_hex(5, 3, 3, periodic=(True,) * 3, jitter=0.2, seed=12, shuffle_seed=13)
    o, g = make_pair(mesh, degree, riemann)
    u = density_wave(o.node_coords(), seed=5)
    g.set_solution_column(u.reshape(-1))
    assert rel(g.get_residual_column(), o.residual(u).reshape(-1)) < TOL[riemann]
    dt = o.time_step(u.reshape(-1), 0.1)
    ref = o.rk3(u.reshape(-1), dt, n_steps=2)
    for s in range(2):
        g.rk3_update(s * dt, dt)
    assert rel(g.get_solution_column(), ref) < TOL[riemann]
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree", [1, 3])
def test_single_trilinear_cell_with_every_boundary_family(degree):
    """One trilinear cell (no interior face) whose six faces carry the six
    boundary families, subsonic ones included."""
    mesh = meshgen.box_hex(1, 1, 1, hi=(0.5, 0.4, 0.3))
    mesh["coords"] = mesh["coords"].copy()
    mesh["coords"][7] += (0.03, -0.02, 0.04)  # the corner (1, 1, 1): the cell is no longer a parallelepiped
    state = conservative(1.2, (60.0, 5.0, -3.0), 1e5)
    bcs = [("subsonic_inlet", [1.2e5, 1.0, 0.0, 0.0, 300.0]), ("subsonic_outlet", [0, 0, 0, 0, 0.9e5]),
           ("inviscid_wall", None), ("smart", state), ("supersonic_inlet", state), ("supersonic_outlet", None)]
    kinds = {"smart": 0, "inviscid_wall": 1, "supersonic_outlet": 2, "supersonic_inlet": 3, "subsonic_inlet": 4,
             "subsonic_outlet": 5}
    o = ol.FrOracle(mesh, q=degree + 1, riemann=1, bcs=[(kinds[k], s if s is not None else [0] * 5) for k, s in bcs])
    g = spatial.FiniteElement(mesh, degree=degree, riemann="hllc", bcs=bcs)
    x = o.node_coords()
    u = conservative(1.2 * (1 + 0.02 * np.sin(7 * x[..., 0])), (60.0, 5.0, -3.0), 1e5 * (1 + 0.01 * x[..., 1]))
    g.set_solution_column(u.reshape(-1))
    assert rel(g.get_residual_column(), o.residual(u).reshape(-1)) < 1e-12
    g.close()


@pytest.mark.gpu
def test_curved_free_stream():
    """A uniform flow stays uniform on trilinear cells (the analytic J* obeys
    the metric identities and the face normals are the J* columns' directions)."""
    mesh = periodic_mesh(seed=3)
    g = spatial.FiniteElement(mesh, degree=3, riemann="roe")
    u = conservative(np.ones(g.sample_points().shape[:-1]), (1.0, 0.5, 0.25), 1.0)
    g.set_solution_column(u.reshape(-1))
    assert np.abs(g.get_residual_column()).max() < 1e-11
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree", [2, 3])
def test_curved_rk3_steps_and_time_step(degree):
    o, g = make_pair(periodic_mesh(seed=5), degree, "roe")
    u = density_wave(o.node_coords(), seed=3)
    g.set_solution_column(u.reshape(-1))
    dt = o.time_step(u.reshape(-1), 0.1)
    assert abs(g.get_time_step(0.1) - dt) <= 1e-13 * dt
    ref = o.rk3(u.reshape(-1), dt, n_steps=5)
    for s in range(5):
        g.rk3_update(s * dt, dt)
    assert rel(g.get_solution_column(), ref) < 1e-12
    g.close()


@pytest.mark.gpu
def test_curved_average_limiter():
    """limiter/average.hpp:11-28 with the per-node weights w det(J) of trilinear cells."""
    o, g = make_pair(periodic_mesh(seed=4), 3, "hllc")
    u = density_wave(o.node_coords(), seed=6, amp=0.3)
    g.set_solution_column(u.reshape(-1))
    g.startup_limiter("average")
    want = o.startup_limiter(u.reshape(-1), "average")
    assert np.abs(g.get_solution_column() - want).max() <= 1e-14 * np.abs(want).max()
    g.close()


def test_curved_mesh_rejected_where_unsupported():
    """Host-only: the mixed form and modal DG refuse trilinear hexahedra."""
    mesh = periodic_mesh()
    with pytest.raises(RuntimeError, match="affine"):
        spatial.FiniteElement(mesh, degree=3, riemann="roe", precision="mixed", host_only=True)
    with pytest.raises(RuntimeError, match="affine"):
        spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", host_only=True)
