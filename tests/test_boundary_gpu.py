"""Boundary conditions on the device against the oracle, through the C ABI.

* Subsonic inlet and outlet (riemann/rotated/euler.hpp:181-280, the families
  of dg/general.hpp:146-183 and fr/general.hpp:566-616): FR and modal-DG
  residuals, and the point-wise fluxes of both stage-kernel forms.
* Boundary values that vary in space and time (the function argument of
  SetSmartBoundary / SetSupersonicInlet, spatial/fem.hpp:196-222): the
  library evaluates the caller's function at the stage times t, t + dt and
  t + dt / 2 of each fused SSP-RK3 step; the oracle's step is driven with the
  same function at the same times (oracle_lib.rk3_timed).  Residuals at a set
  time follow mcfd_set_time (System::SetTime).

Tolerances as the other parity tests: 1e-12 relative to the max norm.
"""
import numpy as np
import pytest

import oracle_lib as ol
from pvcriemannsolvers_amd import meshgen, spatial
from pvcriemannsolvers_amd import riemann as rm

pytestmark = pytest.mark.gpu

GAMMA = 1.4
KIND = {"smart": 0, "inviscid_wall": 1, "supersonic_outlet": 2, "supersonic_inlet": 3, "subsonic_inlet": 4,
        "subsonic_outlet": 5}
RIEMANN = {"exact": 0, "hllc": 1, "roe": 2}


def cons(rho, vel, p):
    rho = np.asarray(rho, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    u = np.zeros(rho.shape + (5,))
    u[..., 0] = rho
    u[..., 1:4] = rho[..., None] * vel
    u[..., 4] = p / (GAMMA - 1) + 0.5 * rho * np.sum(vel * vel, axis=-1)
    return u


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# air-like state: rho ~ 1.2, p ~ 1e5, u ~ 60 along +x (subsonic), T = p / (rho R)
def air_state(x, seed):
    rng = np.random.default_rng(seed)
    shape = x.shape[:-1]
    rho = 1.2 * (1 + 0.02 * np.sin(2 * np.pi * x[..., 0]) + 0.01 * rng.random(shape))
    vel = np.stack([60 + 3 * np.cos(2 * np.pi * x[..., 1]), 2 * rng.random(shape) - 1, np.zeros(shape)], axis=-1)
    p = 1e5 * (1 + 0.01 * np.cos(2 * np.pi * x[..., 0]))
    return cons(rho, vel, p)


SUBSONIC_BCS = [("subsonic_inlet", [1.03e5, 1.0, 0.0, 0.0, 300.0]), ("subsonic_outlet", [0, 0, 0, 0, 0.98e5]),
                ("inviscid_wall", None), ("inviscid_wall", None), ("inviscid_wall", None), ("inviscid_wall", None)]


def oracle_bcs(bcs):
    return [(KIND[k], s if s is not None else [0.0] * 5) for k, s in bcs]


@pytest.mark.parametrize("riemann", ["hllc", "roe"])
@pytest.mark.parametrize("degree", [1, 3])
def test_fr_subsonic_boundaries(riemann, degree):
    mesh = meshgen.box_hex(5, 3, 2, hi=(1.0, 0.6, 0.4), shuffle_seed=4)
    o = ol.FrOracle(mesh, q=degree + 1, riemann=RIEMANN[riemann], bcs=oracle_bcs(SUBSONIC_BCS))
    g = spatial.FiniteElement(mesh, degree=degree, riemann=riemann, bcs=SUBSONIC_BCS)
    u = air_state(o.node_coords(), degree)
    g.set_solution_column(u.reshape(-1))
    assert rel(g.get_residual_column(), o.residual(u).reshape(-1)) < 1e-12
    dt = o.time_step(u.reshape(-1), 0.1)
    ref = o.rk3(u.reshape(-1), dt, n_steps=3)
    for s in range(3):
        g.rk3_update(s * dt, dt)
    assert rel(g.get_solution_column(), ref) < 1e-12
    g.close()


@pytest.mark.parametrize("name", ["tets", "hexes"])
def test_dg_subsonic_boundaries(name):
    mesh = (meshgen.box_tet(4, 3, 2, hi=(1.0, 0.75, 0.5), jitter=0.1, seed=3) if name == "tets"
            else meshgen.box_hex(5, 3, 2, hi=(1.0, 0.6, 0.4), shuffle_seed=6))
    o = ol.DgOracle(mesh, degree=2, riemann=1, bcs=oracle_bcs(SUBSONIC_BCS))
    g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=SUBSONIC_BCS)
    vals = air_state(g.sample_points(), 7)
    g.approximate(vals)
    u = o.approximate(vals)
    assert rel(g.get_solution_column(), u) < 1e-13
    assert rel(g.get_residual_column(), o.residual(u)) < 1e-12
    g.close()


def test_subsonic_point_fluxes_match_oracle():
    rng = np.random.default_rng(8)
    n = 512
    q, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    frames = q * np.sign(np.linalg.det(q))[:, None, None]
    normal = frames[:, 0]
    d = -normal + 0.3 * rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d[(d * normal).sum(1) > -0.3] = -normal[(d * normal).sum(1) > -0.3]
    left = cons(1.2 + 0.1 * rng.random(n), 50 * d + 5 * rng.standard_normal((n, 3)), 1e5 * (1 + 0.05 * rng.random(n)))
    given_in = np.stack([1.03e5 * (1 + 0.01 * rng.random(n)), d[:, 0], d[:, 1], d[:, 2], 300 + rng.random(n)], 1)
    given_out = np.zeros((n, 5))
    given_out[:, 4] = 0.97e5
    for form in ("fr", "dg"):
        s = rm.Rotated("hllc", n, form)
        got = s.get_flux_on_subsonic_inlet(frames, left, given_in)
        want = ol.face_fluxes(1, 3, frames, left, given_in)
        assert (np.abs(got - want).max(1) / np.abs(want).max(1)).max() < 1e-12
        got = s.get_flux_on_subsonic_outlet(frames, left, given_out)
        want = ol.face_fluxes(1, 4, frames, left, given_out)
        assert (np.abs(got - want).max(1) / np.abs(want).max(1)).max() < 1e-13


# ---------------------------------------------------------------- time-dependent boundary values
LEFT, RIGHT = cons(1.0, (0.1, 0.0, 0.0), 1.0), cons(0.125, (0.0, 0.0, 0.0), 0.1)


def pulsing_left(xyz, t):
    """A smart-boundary state that varies over the face and in time."""
    s = np.broadcast_to(LEFT, (len(xyz), 5)).copy()
    s[:, 0] *= 1 + 0.05 * np.sin(40 * t + 10 * xyz[:, 1])
    s[:, 4] *= 1 + 0.05 * np.cos(30 * t + 7 * xyz[:, 2])
    return s


@pytest.mark.parametrize("scheme", ["fr", "dg"])
def test_time_dependent_smart_boundary(scheme):
    bcs = [("smart", LEFT), ("smart", RIGHT)] + [("inviscid_wall", None)] * 4
    constants = np.array([LEFT, RIGHT] + [[0.0] * 5] * 4)
    funcs = {0: pulsing_left}
    if scheme == "fr":
        mesh = meshgen.box_hex(6, 2, 2, hi=(1.0, 0.2, 0.2), shuffle_seed=3)
        o = ol.FrOracle(mesh, q=3, riemann=1, bcs=oracle_bcs(bcs))
        g = spatial.FiniteElement(mesh, degree=2, riemann="hllc", bcs=bcs)
        x = o.node_coords()
        u = np.where((x[..., 0] < 0.5)[..., None], LEFT, RIGHT).reshape(-1)
        g.set_solution_column(u)
        limit = None
    else:
        mesh = meshgen.box_tet(6, 2, 1, hi=(1.0, 0.3, 0.15), jitter=0.1, seed=4)
        o = ol.DgOracle(mesh, degree=2, riemann=1, bcs=oracle_bcs(bcs))
        o.set_limiter(1)
        g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=bcs, limiter="weno")
        vals = np.where((g.sample_points()[..., 0] < 0.5)[..., None], LEFT, RIGHT)
        g.approximate(vals)
        u, _ = o.limit(o.approximate(vals))
        limit = lambda v: o.limit(v)[0]  # noqa: E731
    g.set_boundary_function(0, pulsing_left)
    # the boundary points the function receives are on the x = 0 face
    assert np.abs(g.boundary_points(0)[:, 0]).max() < 1e-14
    # System path: the residual at a set time
    g.set_time(0.013)
    o.set_boundary_states(o.boundary_states_at(0.013, funcs, constants))
    assert rel(g.get_residual_column(), o.residual(u)) < 1e-12
    # fused RK3 steps: the function at t, t + dt, t + dt / 2
    t, dt = 0.0, 2e-3
    for _ in range(4):
        g.rk3_update(t, dt)
        u = ol.rk3_timed(o, u, t, dt, funcs, constants, limit)
        t += dt
    assert rel(g.get_solution_column(), u) < 1e-12
    # removing the function restores the constant state
    g.set_boundary_function(0, None)
    o.set_boundary_states(None)
    g.set_solution_column(u)
    if limit is not None:
        u = limit(u)
    assert rel(g.get_residual_column(), o.residual(u)) < 1e-12
    g.close()


def frozen_left(xyz, t):
    """A smart-boundary state that varies over the face only."""
    return pulsing_left(xyz, 0.02)


@pytest.mark.parametrize("scheme", ["fr", "dg"])
def test_fixed_boundary_values_equal_a_point_function(scheme):
    """mcfd_set_boundary_values (evaluated once) equals the same time-independent
    boundary function evaluated at every stage time, bit for bit, and mixes with
    a timed function on another tag."""
    bcs = [("smart", LEFT), ("smart", RIGHT)] + [("inviscid_wall", None)] * 4
    if scheme == "fr":
        mesh = meshgen.box_hex(6, 2, 2, hi=(1.0, 0.2, 0.2), shuffle_seed=3)
        make = lambda: spatial.FiniteElement(mesh, degree=2, riemann="hllc", bcs=bcs)  # noqa: E731
    else:
        mesh = meshgen.box_tet(6, 2, 1, hi=(1.0, 0.3, 0.15), jitter=0.1, seed=4)
        make = lambda: spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=bcs,  # noqa: E731
                                             limiter="weno")
    a, b = make(), make()
    for g in (a, b):
        vals = np.where((g.sample_points()[..., 0] < 0.5)[..., None], LEFT, RIGHT)
        g.approximate(vals)
        g.set_boundary_function(1, lambda xyz, t: np.broadcast_to(RIGHT * (1 + 0.1 * np.sin(50 * t)), (len(xyz), 5)))
    a.set_boundary_function(0, frozen_left)
    b.set_boundary_values(0, frozen_left(b.boundary_points(0), 0.0))
    t, dt = 0.0, 2e-3
    for _ in range(4):
        a.rk3_update(t, dt)
        b.rk3_update(t, dt)
        t += dt
    assert np.array_equal(a.get_solution_column(), b.get_solution_column())
    a.close()
    b.close()


def test_boundary_values_on_a_tag_without_faces():
    """A tag that no boundary face carries (here: a periodic box) takes a
    function or fixed values without effect."""
    mesh = meshgen.box_hex(3, 3, 3, periodic=(True,) * 3)
    g = spatial.FiniteElement(mesh, degree=1, riemann="hllc", bcs=[("smart", LEFT)])
    g.set_solution_column(np.broadcast_to(LEFT, (g.n_cells, g.n_per_cell, 5)).reshape(-1))
    before = g.get_residual_column()
    assert len(g.boundary_points(0)) == 0
    g.set_boundary_function(0, pulsing_left)
    g.set_boundary_values(0, np.zeros((0, 5)))
    g.rk3_update(0.0, 1e-3)
    g.set_solution_column(np.broadcast_to(LEFT, (g.n_cells, g.n_per_cell, 5)).reshape(-1))
    assert np.array_equal(g.get_residual_column(), before)
    g.close()


# ---------------------------------------------------------------- start-up limiters
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_startup_limiters_match_oracle(degree):
    """limiter/average.hpp and limiter/majority.hpp on the device: the average
    to 1e-14 (a different summation of det J), the majority bit for bit."""
    mesh = meshgen.box_hex(4, 3, 2, hi=(1.0, 0.6, 0.4), shuffle_seed=5)
    bcs = [("inviscid_wall", None)] * 6
    o = ol.FrOracle(mesh, q=degree + 1, riemann=1, bcs=oracle_bcs(bcs))
    g = spatial.FiniteElement(mesh, degree=degree, riemann="hllc", bcs=bcs)
    x = o.node_coords()
    rng = np.random.default_rng(degree)
    u = 1.0 + 0.3 * rng.random(x.shape[:-1] + (5,))
    g.set_solution_column(u.reshape(-1))
    g.startup_limiter("average")
    assert rel(g.get_solution_column(), o.startup_limiter(u.reshape(-1), "average")) < 1e-14
    n = x.shape[1]
    base = rng.random((x.shape[0], 5))
    u = np.repeat(base[:, None, :], n, axis=1)
    for c in range(x.shape[0]):
        idx = rng.choice(n, (n - 1) // 2, replace=False)
        u[c, idx, :] += rng.random((len(idx), 5))
    g.set_solution_column(u.reshape(-1))
    g.startup_limiter("majority")
    np.testing.assert_array_equal(g.get_solution_column(), o.startup_limiter(u.reshape(-1), "majority"))
    if degree >= 2:
        u[0, : n // 2 + 1, 1] = np.arange(n // 2 + 1.0)  # no majority in one component
        g.set_solution_column(u.reshape(-1))
        with pytest.raises(Exception):
            g.startup_limiter("majority")
        np.testing.assert_array_equal(g.get_solution_column(), u.reshape(-1))  # left unchanged
    g.close()


def test_limiter_reconstruct_and_dg_startup_refusal():
    """limiter::Reconstruct once more on the current state (the reference's
    start-up for tetrahedra runs it twice) equals a second oracle pass; the
    nodal start-up limiters refuse modal DG."""
    mesh = meshgen.box_tet(4, 4, 1, hi=(1.0, 1.0, 0.25), jitter=0.1, seed=6)
    o = ol.DgOracle(mesh, degree=2, riemann=1, bcs=[(1, [0] * 5)] * 6)
    o.set_limiter(1)
    g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=[("inviscid_wall", None)] * 6,
                              limiter="weno")
    x = g.sample_points()
    vals = np.where((x[..., 0] < 0.5)[..., None], cons(1.0, (0, 0, 0), 1.0), cons(0.2, (0, 0, 0), 0.1))
    g.approximate(vals)
    u, _ = o.limit(o.approximate(vals))
    g.limiter_reconstruct()
    u2, flags = o.limit(u)
    assert rel(g.get_solution_column(), u2) < 1e-12
    np.testing.assert_array_equal(g.limiter_flags(), flags)
    with pytest.raises(Exception):
        g.startup_limiter("average")
    g.close()
