"""CPU checks that pin the nodal FR oracle (spatial/fr/general.hpp restated,
Gauss-Legendre solution points as in python/element.py:385-437).

* a linear pressure field with constant density and velocity has a linear
  flux; its FR interpolant is exact and equal states give the physical flux
  at every face, so the residual is -(div F) at every solution point, on
  boxes and on sheared (affine) cells, for every Riemann solver;
* uniform flow is preserved on periodic and bounded boxes;
* the quadrature-weighted sum of the residual is the boundary flux integral
  (zero on a periodic box: discrete conservation);
* c_next = 1 corrections equal the right Radau polynomial derivative.
"""
import numpy as np
import pytest

import oracle_lib as ol
from pvcriemannsolvers_amd import meshgen

GAMMA = 1.4
OUTLET, WALL, SMART = 2, 1, 0


def cons(rho, vel, p):
    rho = np.asarray(rho, dtype=np.float64)
    u = np.zeros(rho.shape + (5,))
    u[..., 0] = rho
    for d in range(3):
        u[..., 1 + d] = rho * vel[d]
    u[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return u


def sheared(mesh):
    a = np.array([[1.0, 0.2, 0.1], [0.0, 0.9, 0.3], [0.05, 0.0, 1.1]])
    return dict(mesh, coords=np.ascontiguousarray(mesh["coords"] @ a.T))


MESHES = {
    "box": lambda: meshgen.box_hex(3, 2, 2, hi=(1.0, 0.6, 0.5), shuffle_seed=1),
    "sheared": lambda: sheared(meshgen.box_hex(2, 3, 2, shuffle_seed=2)),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("riemann", [0, 1, 2])
@pytest.mark.parametrize("q", [2, 3, 4])
def test_linear_pressure_closed_form(name, riemann, q):
    o = ol.FrOracle(MESHES[name](), q=q, riemann=riemann, bcs=[(OUTLET, [0] * 5)] * 6)
    x = o.node_coords()
    vel, grad_p = np.array([0.3, -0.2, 0.1]), np.array([0.05, -0.03, 0.02])
    u = cons(np.ones(x.shape[:-1]), vel, 1.0 + (x - 0.5) @ grad_p)
    r = o.residual(u.reshape(-1)).reshape(u.shape)
    div = np.array([0.0, *grad_p, GAMMA / (GAMMA - 1) * vel @ grad_p])
    tol = 1e-6 if riemann == 0 else 1e-11  # exact solver: FindRoot stops at f >= -1e-8
    np.testing.assert_allclose(r, np.broadcast_to(-div, r.shape), rtol=0, atol=tol * np.abs(div).max())


@pytest.mark.parametrize("q", [2, 3, 4])
def test_free_stream_and_periodic_conservation(q):
    mesh = meshgen.box_hex(3, 3, 2, periodic=(True,) * 3, shuffle_seed=5)
    o = ol.FrOracle(mesh, q=q, riemann=2)
    x = o.node_coords()
    u = cons(np.ones(x.shape[:-1]), (0.4, -0.3, 0.2), np.ones(x.shape[:-1]))
    assert np.abs(o.residual(u.reshape(-1))).max() < 1e-12
    rng = np.random.default_rng(q)
    u = cons(1 + 0.2 * np.sin(2 * np.pi * x[..., 0]) + 0.01 * rng.random(x.shape[:-1]), (0.4, -0.3, 0.2),
             1 + 0.1 * np.cos(2 * np.pi * x[..., 1]))
    r = o.residual(u.reshape(-1)).reshape(u.shape)
    w = o.node_weights()
    total = np.einsum("cn,cnk->k", w, r)
    assert np.abs(total).max() < 1e-13


def test_walls_conserve_mass_and_energy():
    mesh = meshgen.box_hex(3, 2, 2, shuffle_seed=3)
    o = ol.FrOracle(mesh, q=3, riemann=1, bcs=[(WALL, [0] * 5)] * 6)
    x = o.node_coords()
    u = cons(1 + 0.1 * np.sin(3 * x[..., 0]), (0.1, 0.2, -0.1), 1 + 0.1 * np.cos(4 * x[..., 2]))
    r = o.residual(u.reshape(-1)).reshape(u.shape)
    total = np.einsum("cn,cnk->k", o.node_weights(), r)
    assert abs(total[0]) < 1e-13 and abs(total[4]) < 1e-13


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_dg_correction_is_right_radau(degree):
    """Vincent(c_next = 1): g_R = (P_k + P_{k+1}) / 2, the right Radau polynomial (Huynh 2007)."""
    from numpy.polynomial import legendre as L
    for s in np.linspace(-1, 1, 7):
        c = np.zeros(degree + 2)
        c[degree] = c[degree + 1] = 0.5
        assert abs(ol.vincent(degree, 1.0, s)[0] - L.legval(s, c)) < 1e-14
        assert abs(ol.vincent(degree, 1.0, s)[2] - L.legval(s, L.legder(c))) < 1e-12
