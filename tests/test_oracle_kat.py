"""CPU checks that pin the oracle against the reference's own known answers.

Every value and tolerance below is transcribed from the reference's tests
(cited per test); the rest are identities the mathematics fixes.  These tests
need no GPU.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_lib as O

GAMMA = 1.4
EXACT, HLLC, ROE = 0, 1, 2
IDENTITY = np.eye(3)


def cons(rho, u, v, w, p):
    return O.primitive_to_conservative([rho, u, v, w, p])


def flux_x(prim):
    """Flux of a primitive state along x (Exact::GetFlux in the normal frame)."""
    return O.flux_matrix(O.primitive_to_conservative(prim))[:, 0]


def expect_near_ratio(actual, expect, eps):
    """test/riemann/euler/exact.cpp:20-28: zero must be exact, else |y/x - 1| <= eps."""
    for a, e in zip(actual, expect):
        if e == 0:
            assert a == 0
        else:
            assert abs(a / e - 1) <= eps, (a, e)


# ---------------------------------------------------------------- gas and flux (test/riemann/euler/types.cpp:55-68)
def test_converters_round_trip():
    rho, u, v, p = 0.1, 0.2, -0.2, 0.3
    c = O.primitive_to_conservative([rho, u, v, 0.0, p])
    assert c[0] == rho and c[1] == rho * u and c[2] == rho * v and c[3] == 0.0
    assert c[4] == p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    back = O.conservative_to_primitive(c)
    np.testing.assert_allclose(back, [rho, u, v, 0.0, p], rtol=4 * np.finfo(float).eps)


def test_nonphysical_states_are_zeroed():
    # riemann/euler/types.hpp:195-202: negative density or pressure becomes zero.
    pv = O.conservative_to_primitive([1.0, 3.0, 0.0, 0.0, 1.0])  # E < kinetic energy
    assert pv[4] == 0.0
    pv = O.conservative_to_primitive([-1.0, 0.0, 0.0, 0.0, 1.0])
    assert pv[0] == 0.0


def test_flux_matrix_closed_form():
    rho, u, v, w, p = 1.3, 0.4, -0.7, 0.25, 2.5
    f = O.flux_matrix(cons(rho, u, v, w, p))
    vel = np.array([u, v, w])
    e = p / (GAMMA - 1) + 0.5 * rho * vel @ vel
    for d in range(3):
        ref = [rho * vel[d], rho * u * vel[d], rho * v * vel[d], rho * w * vel[d], (e + p) * vel[d]]
        ref[1 + d] += p
        np.testing.assert_allclose(f[:, d], ref, rtol=1e-15, atol=1e-15)


# ---------------------------------------------------------------- exact solver (test/riemann/euler/exact.cpp)
# (left, right, expected t-axis state) in (rho, u, p); transverse velocities as in TestExact3d.
EXACT_CASES = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), (0.426319, +0.927453, 0.303130), "left"),     # :51-57
    ((0.125, 0.0, 0.1), (1.0, 0.0, 1.0), (0.426319, -0.927453, 0.303130), "right"),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), (5.99924, 19.5975, 460.894), "left"),  # :261-266
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), (0.575062, 19.59745, 460.8938), "left"),    # :267-272
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), (0.575113, -6.196328, 46.09504), "right"),   # :273-278
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), (0.21852, 0.0, 0.001894), "left"),             # :279-285
    ((1.0, -4.0, 0.4), (1.0, 4.0, 0.4), (0.0, 0.0, 0.0), "left"),                      # :287-293
]


@pytest.mark.parametrize("left,right,star,side", EXACT_CASES)
def test_exact_known_answers(left, right, star, side):
    """test/riemann/euler/exact.cpp:243-293 (3-D), relative 1e-4 per component."""
    vl, wl, vr, wr = 1.5, 1.5, 2.5, 0.5
    pl = [left[0], left[1], vl, wl, left[2]]
    pr = [right[0], right[1], vr, wr, right[2]]
    f = O.riemann_flux(EXACT, IDENTITY, O.primitive_to_conservative(pl), O.primitive_to_conservative(pr))
    v, w = (vl, wl) if side == "left" else (vr, wr)
    expect = flux_x([star[0], star[1], v, w, star[2]])
    if star[0] == 0.0:  # vacuum: GetFlux of a zero state is exactly zero
        assert np.all(f == 0)
    elif star[1] == 0.0:  # almost vacuum: u* = 0 exactly, mass/energy fluxes vanish
        assert abs(f[0]) < 1e-12 and abs(f[4]) < 1e-12
        assert abs(f[1] / expect[1] - 1) <= 1e-4
    else:
        expect_near_ratio(f[[0, 1, 4]], expect[[0, 1, 4]], 1e-4)


def test_exact_equal_states_is_consistent():
    """test/riemann/euler/exact.cpp:41-50."""
    for u in (0.0, 1.0, -1.0):
        c = cons(1.0, u, 1.5, 1.5, 1.0)
        f = O.riemann_flux(EXACT, IDENTITY, c, c)
        np.testing.assert_allclose(f, flux_x([1.0, u, 1.5, 1.5, 1.0]), rtol=1e-12, atol=1e-14)


def test_exact_left_vacuum_fan_closed_form():
    """test/riemann/euler/exact.cpp:294-346: right state next to vacuum, c_R = 1.

    The t-axis state inside the right rarefaction fan is rho = 1.4*6^-5,
    u = -1/6, p = 6^-7 when u_R = 4.
    """
    right = [1.4, 4.0, 50.0, 20.0, 1.0]
    pv = O.exact_primitive([0.0, -100.0, -50.0, -20.0, 0.0], right)
    assert abs(pv[0] - 1.4 * 6.0 ** -5) <= 1e-14
    assert abs(pv[1] + 1 / 6) <= 1e-14
    assert abs(pv[4] - 6.0 ** -7) <= 1e-14
    # fan entirely to the right of the axis: the vacuum state
    right[1] = 6.0
    pv = O.exact_primitive([0.0, -100.0, -50.0, -20.0, 0.0], right)
    assert pv[0] == 0 and pv[4] == 0
    # fan entirely to the left: the right state itself
    right[1] = -2.0
    pv = O.exact_primitive([0.0, -100.0, -50.0, -20.0, 0.0], right)
    np.testing.assert_allclose(pv[[0, 1, 4]], [1.4, -2.0, 1.0], rtol=1e-14)


def test_exact_right_vacuum_fan_closed_form():
    """test/riemann/euler/exact.cpp:95-141 (mirror of the left case)."""
    left = [1.4, -4.0, 0.0, 0.0, 1.0]
    pv = O.exact_primitive(left, [0.0, 100.0, 0.0, 0.0, 0.0])
    assert abs(pv[0] - 1.4 * 6.0 ** -5) <= 1e-14
    assert abs(pv[1] - 1 / 6) <= 1e-14
    assert abs(pv[4] - 6.0 ** -7) <= 1e-14


# ---------------------------------------------------------------- HLLC (test/riemann/euler/hllc.cpp)
def ratio(x, y):
    return abs(x - y) / max(abs(x), abs(y))


def test_hllc_consistency():
    """hllc.cpp:40-49: F(u, u) = F(u) within 1e-16 (norm of the difference)."""
    rng = np.random.default_rng(31415926)
    for _ in range(64):
        rho, u, p = rng.random(3) + 1e-3
        c = cons(rho, u, rng.random(), rng.random(), p)
        f = O.riemann_flux(HLLC, IDENTITY, c, c)
        assert np.linalg.norm(f - O.flux_matrix(c)[:, 0]) <= 1e-15 * max(1.0, np.abs(f).max())


@pytest.mark.parametrize("left,right,star,side", EXACT_CASES[:5])
def test_hllc_close_to_exact(left, right, star, side):
    """hllc.cpp:50-75 and the 3-D block: mass and energy within 5 %, momentum within 25 %."""
    pl = [left[0], left[1], 1.5, 3.5, left[2]]
    pr = [right[0], right[1], 2.5, 4.5, right[2]]
    f = O.riemann_flux(HLLC, IDENTITY, O.primitive_to_conservative(pl), O.primitive_to_conservative(pr))
    v, w = (1.5, 3.5) if side == "left" else (2.5, 4.5)
    e = flux_x([star[0], star[1], v, w, star[2]])
    assert ratio(f[0], e[0]) <= 0.05 and ratio(f[4], e[4]) <= 0.05 and ratio(f[1], e[1]) <= 0.25


# ---------------------------------------------------------------- Roe (textbook; parity of values unpinned)
def test_roe_consistency_and_mirror_symmetry():
    """Roe flux: F(u, u) = F(u); swapping sides and reflecting the normal velocity reflects the flux."""
    rng = np.random.default_rng(7)
    for _ in range(64):
        a = cons(0.5 + rng.random(), rng.normal(), rng.normal(), rng.normal(), 0.5 + rng.random())
        b = cons(0.5 + rng.random(), rng.normal(), rng.normal(), rng.normal(), 0.5 + rng.random())
        np.testing.assert_allclose(O.riemann_flux(ROE, IDENTITY, a, a), O.flux_matrix(a)[:, 0],
                                   rtol=1e-13, atol=1e-13)
        f = O.riemann_flux(ROE, IDENTITY, a, b)
        ra, rb = a.copy(), b.copy()
        ra[1] *= -1
        rb[1] *= -1
        g = O.riemann_flux(ROE, IDENTITY, rb, ra)
        np.testing.assert_allclose(g * [-1, 1, -1, -1, -1], f, rtol=1e-12, atol=1e-12)


def test_roe_exact_on_isolated_shock_and_contact():
    """A Rankine-Hugoniot shock or a contact is resolved exactly by Roe's linearisation (Toro, sec. 11.2)."""
    # contact: equal u and p, different rho -> flux is upwind state flux
    l, r = cons(1.0, 0.3, 0.1, 0.0, 1.0), cons(0.2, 0.3, -0.4, 0.0, 1.0)
    np.testing.assert_allclose(O.riemann_flux(ROE, IDENTITY, l, r), O.flux_matrix(l)[:, 0], rtol=1e-12, atol=1e-13)
    # stationary normal shock, Mach 2 upstream
    m, g = 2.0, GAMMA
    rho1, p1 = 1.0, 1.0
    u1 = m * math.sqrt(g * p1 / rho1)
    rho2 = rho1 * (g + 1) * m * m / ((g - 1) * m * m + 2)
    p2 = p1 * (2 * g * m * m - (g - 1)) / (g + 1)
    u2 = rho1 * u1 / rho2
    l, r = cons(rho1, u1, 0, 0, p1), cons(rho2, u2, 0, 0, p2)
    f = O.riemann_flux(ROE, IDENTITY, l, r)
    np.testing.assert_allclose(f, O.flux_matrix(l)[:, 0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(f, O.flux_matrix(r)[:, 0], rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- rotated frame (test/riemann/rotated/euler.cpp)
@pytest.mark.parametrize("kind", [EXACT, HLLC, ROE])
def test_rotational_invariance(kind):
    """rotated/euler.cpp:94-117: frames (x, y, z) and (x, -y, -z) give the same Sod flux."""
    l = cons(1.0, 0.0, 1.5, 1.5, 1.0)
    r = cons(0.125, 0.0, 2.5, 0.5, 0.1)
    f1 = O.riemann_flux(kind, IDENTITY, l, r)
    f2 = O.riemann_flux(kind, np.diag([1.0, -1.0, -1.0]), l, r)
    np.testing.assert_allclose(f1, f2, rtol=1e-14, atol=1e-15)
    # a random rotation R: F_R(R u) = R F(u) for the momentum rows
    rng = np.random.default_rng(kind)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 2] *= -1
    frame = q  # rows are the old axes expressed in the rotated coordinates

    def rot(c):
        out = c.copy()
        out[1:4] = q.T @ c[1:4]
        return out

    fr = O.riemann_flux(kind, frame, rot(l), rot(r))
    ref = O.riemann_flux(kind, IDENTITY, l, r)
    np.testing.assert_allclose(fr[[0, 4]], ref[[0, 4]], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(fr[1:4], q.T @ ref[1:4], rtol=1e-12, atol=1e-12)


def test_inviscid_wall_flux():
    """rotated/euler.cpp:119-150: wall flux = (0, p n, 0) with exact zeros."""
    rng = np.random.default_rng(31415926)
    for _ in range(256):
        ang = math.radians(180 * rng.uniform(-1, 1))
        c, s = math.cos(ang), math.sin(ang)
        frame = np.array([[1, 0, 0], [0, c, s], [0, -s, c]], dtype=float)
        rho = 1.29 * (1 + 0.05 * rng.uniform(-1, 1))
        p = 101325 * (1 + 0.05 * rng.uniform(-1, 1))
        st = cons(rho, 10 * rng.uniform(-1, 1), 20 * rng.uniform(-1, 1), 30 * rng.uniform(-1, 1), p)
        f = O.wall_flux(frame, st)
        assert abs(f[1] - p) <= 1e-10 * p
        assert f[0] == 0.0 and f[2] == 0.0 and f[3] == 0.0 and f[4] == 0.0


# ---------------------------------------------------------------- quadrature and bases
@pytest.mark.parametrize("lobatto,q", [(False, q) for q in range(1, 7)] + [(True, q) for q in range(2, 7)])
def test_quadrature_monomial_exactness(lobatto, q):
    """test/integrator/legendre.cpp, lobatto.cpp: exact up to degree 2q-1 (2q-3 for Lobatto)."""
    x, w = O.quadrature(lobatto, q)
    top = 2 * q - 3 if lobatto else 2 * q - 1
    for k in range(top + 1):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) <= (1e-14 if q >= 5 else 1e-15)
    if lobatto:
        assert x[0] == -1.0 and x[-1] == 1.0
    np.testing.assert_allclose(x, -x[::-1], atol=1e-15)


@pytest.mark.parametrize("q", [2, 3, 4, 5])
def test_lagrange_basis_identities(q):
    """test/basis/lagrange.cpp:34-160: Kronecker delta, partition of unity, derivative sum zero."""
    x, _ = O.quadrature(False, q)
    for i in range(q):
        v, d = O.lagrange(x, x[i])
        np.testing.assert_allclose(v, np.eye(q)[i], atol=1e-15)
    rng = np.random.default_rng(q)
    for s in rng.uniform(-1, 1, 16):
        v, d = O.lagrange(x, s)
        assert abs(v.sum() - 1) <= 1e-14 and abs(d.sum()) <= 1e-12
        # reproduces x^k and its derivative for k < q
        for k in range(q):
            assert abs(v @ x ** k - s ** k) <= 1e-13
            assert abs(d @ x ** k - (k * s ** (k - 1) if k else 0.0)) <= 1e-12


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("lumping", [False, True])
def test_vincent_correction(degree, lumping):
    """test/basis/vincent.cpp:24-81: g_R(1)=1, g_R(-1)=0, orthogonality, g_R'(s) = -g_L'(-s).

    c_next = 1 is Vincent::DiscontinuousGalerkin (orthogonal to P_{degree-1});
    c_next = degree/(2 degree + 1) is HuynhLumpingLobatto (orthogonal to P_{degree-2},
    g_R'(-1) = 0).
    """
    c_next = degree / (2.0 * degree + 1) if lumping else 1.0
    g = O.vincent(degree, c_next, 1.0)
    assert g[0] == 1.0 and abs(g[1]) <= 1e-16
    g = O.vincent(degree, c_next, -1.0)
    assert abs(g[0]) <= 1e-16 and g[1] == 1.0
    x, w = O.quadrature(False, 6)
    gr = np.array([O.vincent(degree, c_next, s)[0] for s in x])
    for ell in range(degree - 1 if lumping else degree):
        p = np.polynomial.legendre.legval(x, np.eye(ell + 1)[ell])
        assert abs(np.sum(w * gr * p)) <= 1e-15
    rng = np.random.default_rng(31415926)
    for s in rng.uniform(-1, 1, 64):
        a, b = O.vincent(degree, c_next, s), O.vincent(degree, c_next, -s)
        assert a[2] == -b[3]
        approx = (O.vincent(degree, c_next, s + 1e-6)[1] - O.vincent(degree, c_next, s - 1e-6)[1]) / 2e-6
        assert abs(a[3] / approx - 1) <= 1e-6
    if lumping:
        assert abs(O.vincent(degree, c_next, -1.0)[2]) <= 1e-15
        assert abs(O.vincent(degree, c_next, 1.0)[3]) <= 1e-15
