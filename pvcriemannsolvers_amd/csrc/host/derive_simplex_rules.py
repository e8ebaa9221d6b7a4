"""Derive the symmetric simplex quadrature rules used by the tetrahedral DG path.

The rules are fixed by their symmetry orbits and degree of exactness:

* tetrahedron, 14 points, degree 5: two S31 orbits (a, a, a, 1-3a) and one
  S22 orbit (b, b, 1/2-b, 1/2-b);
* triangle, 6 points, degree 4: two S21 orbits (a, a, 1-2a).

This script solves the moment equations (every monomial up to the degree
integrated exactly over the reference simplex) by Gauss-Newton in 50-digit
arithmetic and prints the orbit parameters and weights with 20 significant
digits.  The printed values are the constants of simplex_rules.hpp; the
reference's integrator/tetrahedron.hpp:150-187 and integrator/triangle.hpp:
163-188 use rules with the same orbit structure and point order.
"""
from math import factorial

import mpmath as mp

mp.mp.dps = 50


def tet_points(a1, a2, b):
    pts = []
    for a in (a1, a2):
        c = 1 - 3 * a
        pts += [(a, a, a), (a, a, c), (a, c, a), (c, a, a)]
    c = (1 - 2 * b) / 2
    pts += [(b, b, c), (b, c, b), (b, c, c), (c, b, b), (c, b, c), (c, c, b)]
    return pts


def tri_points(a1, a2):
    pts = []
    for a in (a1, a2):
        c = 1 - 2 * a
        pts += [(a, a), (a, c), (c, a)]
    return pts


def tet_residual(v):
    a1, a2, b, w1, w2, w3 = v
    pts = tet_points(a1, a2, b)
    ws = [w1] * 4 + [w2] * 4 + [w3] * 6
    res = []
    for i in range(6):
        for j in range(6 - i):
            for k in range(6 - i - j):
                exact = mp.mpf(factorial(i) * factorial(j) * factorial(k)) / factorial(i + j + k + 3)
                res.append(mp.fsum(w * p[0] ** i * p[1] ** j * p[2] ** k for w, p in zip(ws, pts)) - exact)
    return res


def tri_residual(v):
    a1, a2, w1, w2 = v
    pts = tri_points(a1, a2)
    ws = [w1] * 3 + [w2] * 3
    res = []
    for i in range(5):
        for j in range(5 - i):
            exact = mp.mpf(factorial(i) * factorial(j)) / factorial(i + j + 2)
            res.append(mp.fsum(w * p[0] ** i * p[1] ** j for w, p in zip(ws, pts)) - exact)
    return res


def gauss_newton(f, x0, iters=60):
    x = mp.matrix([mp.mpf(v) for v in x0])
    for _ in range(iters):
        r = mp.matrix(f(list(x)))
        jac = mp.matrix(len(r), len(x))
        for c in range(len(x)):
            h = mp.mpf(10) ** -30
            xp = x.copy()
            xp[c] += h
            rp = mp.matrix(f(list(xp)))
            for rr in range(len(r)):
                jac[rr, c] = (rp[rr] - r[rr]) / h
        step = mp.lu_solve(jac.T * jac, jac.T * r)
        x -= step
        if mp.norm(step) < mp.mpf(10) ** -45:
            break
    return list(x), mp.norm(mp.matrix(f(list(x))))


if __name__ == "__main__":
    tet, tet_res = gauss_newton(tet_residual, [0.31, 0.09, 0.045, 0.019, 0.012, 0.007])
    tri, tri_res = gauss_newton(tri_residual, [0.45, 0.09, 0.11, 0.055])
    print("tetrahedron 14 (a1, a2, b, w1, w2, w3), residual", mp.nstr(tet_res, 3))
    for v in tet:
        print("  ", mp.nstr(v, 20))
    print("triangle 6 (a1, a2, w1, w2), residual", mp.nstr(tri_res, 3))
    for v in tri:
        print("  ", mp.nstr(v, 20))
