"""The Python examples of the integration guide, run on a small mesh for a few steps.

The guide drives a limited modal-DG handle on jittered tetrahedra with
`rk3_update` and states that `temporal.RungeKutta(3).update(fe, t, dt)` gives
the same columns; it then sets a boundary function, writes a frame and a
restart file and reads the restart back.
"""
import os

import numpy as np
import pytest

from pvcriemannsolvers_amd import meshgen, restart, spatial, temporal, vtk

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def cons(rho, vel, p):
    rho = np.asarray(rho, dtype=np.float64)
    u = np.zeros(rho.shape + (5,))
    u[..., 0] = rho
    for d in range(3):
        u[..., 1 + d] = rho * vel[d]
    u[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return u


def values_at(x, t=0.0):
    """Four constant quadrants of the unit square (the 2-D Riemann problem of riemann2d.cpp), seeded; any t."""
    rng = np.random.default_rng(21)
    rho, p = rng.uniform(0.1, 1.5, 4), rng.uniform(0.03, 1.5, 4)
    vel = rng.uniform(-1, 1, (4, 2))
    q = (x[..., 0] > 0.5).astype(int) + 2 * (x[..., 1] > 0.5).astype(int)
    return cons(rho[q], (vel[q, 0], vel[q, 1], 0.0 * q), p[q])


def make(bcs=(("inviscid_wall", None),) * 6):
    mesh = meshgen.box_tet(8, 8, 2, hi=(1.0, 1.0, 0.25), jitter=0.15, seed=5)
    return spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=list(bcs), limiter="weno")


def test_driver_loop_and_runge_kutta_give_the_same_columns():
    fe, twin = make(), make()
    for h in (fe, twin):
        h.approximate(values_at(h.sample_points()))
    np.testing.assert_array_equal(fe.get_solution_column(), twin.get_solution_column())
    rk = temporal.RungeKutta(3)
    t, t_end = 0.0, 1.0
    for _ in range(3):
        dt = fe.get_time_step(0.1, t_end - t)
        assert dt == twin.get_time_step(0.1, t_end - t)
        fe.rk3_update(t, dt)
        rk.update(twin, t, dt)
        t += dt
    assert fe.limiter_flags().any()
    u = fe.get_solution_column()
    assert np.isfinite(u).all()
    np.testing.assert_array_equal(u, twin.get_solution_column())
    fe.close()
    twin.close()


def test_boundary_function_frame_and_restart(tmp_path):
    fe = make(bcs=[("smart", cons(1.0, (0.0, 0.0, 0.0), 1.0))] + [("inviscid_wall", None)] * 5)
    fe.approximate(values_at(fe.sample_points()))
    fe.set_boundary_function(0, lambda xyz, t: values_at(xyz, t))
    t = 0.0
    for _ in range(2):
        dt = fe.get_time_step(0.1)
        fe.rk3_update(t, dt)
        t += dt
    out = str(tmp_path / "out")
    vtk.write_solutions(fe, out, "frame_0010")
    assert os.path.isfile(os.path.join(out, "frame_0010.pvtu"))
    piece = vtk.read_piece(os.path.join(out, "frame_0010", "0.vtu"))
    assert len(piece["offsets"]) == fe.n_cells
    assert np.isfinite(piece["point_data"]["Density"]).all()
    path = str(tmp_path / "restart.npz")
    restart.write_restart(fe, path, t)
    u = fe.get_solution_column()
    fe.rk3_update(t, fe.get_time_step(0.1))  # move on, then come back
    assert not np.array_equal(fe.get_solution_column(), u)
    # restored as stored, with no limiter pass (the reference's restart path)
    assert restart.read_restart(fe, path) == t
    np.testing.assert_array_equal(fe.get_solution_column(), u)
    fe.close()
