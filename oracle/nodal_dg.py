"""Oracle (test infrastructure): collocated nodal DG on hexahedra in weak form.

This is a second, independently written CPU statement of the hexahedral
residual.  It is used only by tests/ to pin the FR restatement in
oracle/src/fr.hpp; nothing in the product imports it.

Scheme: the arithmetic of dg::General (spatial/dg/general.hpp:58-69 volume
term sum_q W_q F(u_q) . grad(phi_i)(x_q), :70-89 face term -/+ W_f F^ phi_i,
boundary families :92-202) with a tensor-product Lagrange basis on the Q^3
solution points, quadrature on those same points (collocation), and the
diagonal mass division of dg/lobatto.hpp:143-158 (1 / (w_q |det J_q|)).

What is shared with the FR restatement, and only that:
  * the node order ijk = (i Q + j) Q + k of polynomial/hexahedron.hpp
    (basis/lagrange.hpp:239-241), so the two columns can be compared;
  * the upwind flux, passed in as a callable (the Riemann solvers are pinned
    on their own by the reference's known answers).
Everything else is computed here from the mesh arrays alone: 1-D rules from
numpy, Lagrange values from their product formula, the trilinear map and its
Jacobian from the HEXA_8 corners, face pairing from vertex sets (periodic
faces by translated centres), face points matched by physical position,
normals and area elements from cross products of the face tangents.

Known equivalences that make this a pin (Huynh 2007, "A flux reconstruction
approach to high-order schemes including discontinuous Galerkin methods"):
  * Gauss-Legendre points: FR with the DG correction (Radau, c = 0 in
    Vincent's family, c_next = 1 here) equals this collocated nodal DG;
  * Gauss-Lobatto points: FR with Huynh's lumped-Lobatto correction
    (c_next = p / (2p + 1), fr/lobatto.hpp) equals this scheme (DGSEM, the
    arithmetic of dg::Lobatto, dg/lobatto.hpp:58-191).
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as npleg

GAMMA = 1.4

# HEXA_8 corners in reference coordinates (CGNS order).
CORNERS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)

# Boundary kinds of the BC table (same meaning as the product's table).
SMART, WALL, OUTLET, INLET, SUB_INLET, SUB_OUTLET = 0, 1, 2, 3, 4, 5


def rule(q: int, lobatto: bool):
    """1-D Gauss-Legendre or Gauss-Lobatto points and weights on [-1, 1]."""
    if not lobatto:
        return npleg.leggauss(q)
    # interior points: roots of P'_{q-1}; weights 2 / (q (q - 1) P_{q-1}(x)^2)
    c = np.zeros(q)
    c[q - 1] = 1.0
    inner = np.sort(npleg.legroots(npleg.legder(c))) if q > 2 else np.zeros(0)
    x = np.concatenate([[-1.0], inner, [1.0]])
    w = 2.0 / (q * (q - 1) * npleg.legval(x, c) ** 2)
    return x, w


def lagrange(x: np.ndarray, s: np.ndarray):
    """Values L[t, m] = l_m(s_t) and derivatives dL[t, m] = l_m'(s_t)."""
    s = np.atleast_1d(np.asarray(s, dtype=np.float64))
    q = len(x)
    val = np.ones((len(s), q))
    der = np.zeros((len(s), q))
    for m in range(q):
        others = [j for j in range(q) if j != m]
        denom = np.prod([x[m] - x[j] for j in others])
        for j in others:
            val[:, m] *= (s - x[j])
        val[:, m] /= denom
        for jd in others:  # product rule
            term = np.ones(len(s))
            for j in others:
                if j != jd:
                    term *= (s - x[j])
            der[:, m] += term
        der[:, m] /= denom
    return val, der


def trilinear(x8: np.ndarray, xi: np.ndarray):
    """Points x(xi) [n, 3] and Jacobians J[n, d, e] = dx_d / dxi_e of a HEXA_8 cell."""
    xi = np.atleast_2d(xi)
    f = 1.0 + xi[:, None, :] * CORNERS[None, :, :]          # [n, 8, 3]
    n = np.prod(f, axis=2) / 8.0                             # shape functions
    dn = np.empty(f.shape)
    for e in range(3):
        other = [d for d in range(3) if d != e]
        dn[:, :, e] = CORNERS[None, :, e] * f[:, :, other[0]] * f[:, :, other[1]] / 8.0
    return n @ x8, np.einsum("vd,nve->nde", x8, dn)


def euler_flux(u: np.ndarray) -> np.ndarray:
    """Physical flux F[..., k, d] of conservative states u[..., 5]."""
    rho = u[..., 0]
    vel = u[..., 1:4] / rho[..., None]
    p = (GAMMA - 1.0) * (u[..., 4] - 0.5 * rho * np.sum(vel * vel, axis=-1))
    f = np.empty(u.shape[:-1] + (5, 3))
    f[..., 0, :] = u[..., 1:4]
    f[..., 1:4, :] = u[..., 1:4, None] * vel[..., None, :]
    for d in range(3):
        f[..., 1 + d, d] += p
    f[..., 4, :] = (u[..., 4] + p)[..., None] * vel
    return f


def pressure(u: np.ndarray):
    rho = u[..., 0]
    return (GAMMA - 1.0) * (u[..., 4] - 0.5 * np.sum(u[..., 1:4] ** 2, axis=-1) / rho)


def _frame(n: np.ndarray) -> np.ndarray:
    """Rows (n, t, b) of an orthonormal frame whose first row is n."""
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t = a - (a @ n) * n
    t /= np.linalg.norm(t)
    return np.stack([n, t, np.cross(n, t)])


class NodalDgWeak:
    """Residual du/dt of collocated nodal DG on a hexahedral mesh (mesh dict of
    pvcriemannsolvers_amd.meshgen).  ``bcs`` is a list of (kind, conservative
    state) indexed by boundary tag; negative tags are periodic faces."""

    def __init__(self, mesh: dict, q: int, bcs=(), lobatto: bool = False):
        self.q = q
        self.n = q ** 3
        self.coords = np.asarray(mesh["coords"], dtype=np.float64)
        self.cells = np.asarray(mesh["cells"], dtype=np.int64)
        self.bcs = list(bcs)
        self.x1, self.w1 = rule(q, lobatto)
        l_nodes, d_nodes = lagrange(self.x1, self.x1)
        self.l_nodes = l_nodes  # identity up to rounding
        ii, jj, kk = np.meshgrid(np.arange(q), np.arange(q), np.arange(q), indexing="ij")
        self.node_idx = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)  # node -> (i, j, k)
        self.xi_nodes = self.x1[self.node_idx]
        self.w_nodes = np.prod(self.w1[self.node_idx], axis=1)
        # reference gradients G[e][q_node, i_node] = d phi_i / d xi_e at node q
        eye = np.eye(q)
        mats = []
        for e in range(3):
            factors = [eye, eye, eye]
            factors[e] = d_nodes  # d_nodes[t, m] = l_m'(x_t)
            mats.append(np.kron(np.kron(factors[0], factors[1]), factors[2]))
        self.grad_ref = np.stack(mats)  # [3, n, n]
        nc = len(self.cells)
        self.x_nodes = np.empty((nc, self.n, 3))
        self.jac_nodes = np.empty((nc, self.n, 3, 3))
        for c in range(nc):
            self.x_nodes[c], self.jac_nodes[c] = trilinear(self.coords[self.cells[c]], self.xi_nodes)
        self.det_nodes = np.linalg.det(self.jac_nodes)
        self._build_faces(mesh)

    # ---- connectivity, from vertex sets alone
    @staticmethod
    def _ref_faces():
        """(direction e, side s) of the six reference faces."""
        return [(e, s) for e in range(3) for s in (-1, 1)]

    def _face_corners(self, c, e, s):
        on = np.nonzero(CORNERS[:, e] == s)[0]
        return self.cells[c][on]

    def _build_faces(self, mesh):
        bnodes = np.asarray(mesh["bface_nodes"], dtype=np.int64)
        btags = np.asarray(mesh["bface_tags"], dtype=np.int64)
        period = np.asarray(mesh["period"], dtype=np.float64)
        tag_of = {tuple(sorted(f)): int(t) for f, t in zip(bnodes, btags)}
        by_key = {}
        for c in range(len(self.cells)):
            for e, s in self._ref_faces():
                key = tuple(sorted(self._face_corners(c, e, s)))
                by_key.setdefault(key, []).append((c, e, s))
        self.interior = []   # (cell A, e, s, cell B, e, s, shift B - A)
        self.boundary = []   # (cell, e, s, tag)
        periodic = []
        for key, sides in by_key.items():
            if len(sides) == 2:
                self.interior.append((*sides[0], *sides[1], np.zeros(3)))
                continue
            assert len(sides) == 1, "non-manifold face"
            tag = tag_of[key]
            if tag >= 0:
                self.boundary.append((*sides[0], tag))
            else:
                periodic.append((sides[0], self.coords[list(key)].mean(axis=0)))
        # periodic pairs: centres differ by one period vector
        used = [False] * len(periodic)
        scale = max(1.0, np.abs(self.coords).max())
        for a, (sa, ca) in enumerate(periodic):
            if used[a]:
                continue
            for b in range(a + 1, len(periodic)):
                if used[b]:
                    continue
                sb, cb = periodic[b]
                delta = cb - ca
                for d in range(3):
                    if period[d] <= 0:
                        continue
                    shift = np.zeros(3)
                    shift[d] = np.sign(delta[d]) * period[d]
                    if np.linalg.norm(delta - shift) < 1e-10 * scale:
                        self.interior.append((*sa, *sb, shift))
                        used[a] = used[b] = True
                        break
                if used[a]:
                    break
            assert used[a], "unpaired periodic face"

    # ---- face geometry
    def _face_points(self, c, e, s):
        """Reference points, physical points, weights (w_a w_b dS) and outward
        unit normals of the Q x Q points of face (e, s) of cell c."""
        other = [d for d in range(3) if d != e]
        ga, gb = np.meshgrid(np.arange(self.q), np.arange(self.q), indexing="ij")
        ga, gb = ga.ravel(), gb.ravel()
        xi = np.empty((len(ga), 3))
        xi[:, e] = s
        xi[:, other[0]] = self.x1[ga]
        xi[:, other[1]] = self.x1[gb]
        x, jac = trilinear(self.coords[self.cells[c]], xi)
        area = np.cross(jac[:, :, other[0]], jac[:, :, other[1]])
        ds = np.linalg.norm(area, axis=1)
        normal = area / ds[:, None]
        # outward: along +grad(xi_e) for s = +1
        grad_e = np.linalg.inv(jac)[:, e, :]
        sign = np.sign(np.einsum("nd,nd->n", normal, grad_e)) * s
        normal *= sign[:, None]
        return xi, x, self.w1[ga] * self.w1[gb] * ds, normal

    def _basis(self, xi):
        """phi_i(xi) for every node i: [n_points, n]."""
        vals = [lagrange(self.x1, xi[:, d])[0] for d in range(3)]
        idx = self.node_idx
        return vals[0][:, idx[:, 0]] * vals[1][:, idx[:, 1]] * vals[2][:, idx[:, 2]]

    # ---- residual
    def residual(self, u: np.ndarray, upwind, boundary=None) -> np.ndarray:
        """du/dt in the column layout [cell][node][5].  ``upwind(frame, ul, ur)``
        is the rotated upwind flux with frame rows (n, t, b), n from ul to ur;
        ``boundary(kind, frame, u_in, given)`` the subsonic inlet and outlet
        fluxes, whose formulas are pinned on their own."""
        nc = len(self.cells)
        u = np.asarray(u, dtype=np.float64).reshape(nc, self.n, 5)
        r = np.zeros_like(u)
        # volume: sum_q W_q F(u_q) . grad_x phi_i(x_q)
        f = euler_flux(u)                                              # [c, q, k, d]
        jinv = np.linalg.inv(self.jac_nodes)                           # [c, q, e, d]
        wq = self.w_nodes[None, :] * np.abs(self.det_nodes)            # [c, q]
        # grad_x phi_i(q)_d = sum_e jinv[e, d] dphi_i/dxi_e(q)
        flux_ref = np.einsum("cqkd,cqed->cqke", f, jinv) * wq[:, :, None, None]
        r += np.einsum("cqke,eqi->cik", flux_ref, self.grad_ref)
        # interior faces
        for ca, ea, sa, cb, eb, sb, shift in self.interior:
            xi_a, x_a, wf, n_a = self._face_points(ca, ea, sa)
            xi_b_all, x_b_all, _, _ = self._face_points(cb, eb, sb)
            match = np.empty(len(x_a), dtype=np.int64)
            h = np.linalg.norm(x_a.max(axis=0) - x_a.min(axis=0))
            for p in range(len(x_a)):
                dist = np.linalg.norm(x_b_all - (x_a[p] + shift), axis=1)
                match[p] = int(np.argmin(dist))
                assert dist[match[p]] < 1e-9 * max(h, 1.0), "face points do not coincide"
            xi_b = xi_b_all[match]
            phi_a, phi_b = self._basis(xi_a), self._basis(xi_b)
            ua, ub = phi_a @ u[ca], phi_b @ u[cb]
            for p in range(len(x_a)):
                fl = upwind(_frame(n_a[p]), ua[p], ub[p]) * wf[p]
                r[ca] -= np.outer(phi_a[p], fl)
                r[cb] += np.outer(phi_b[p], fl)
        # boundary faces
        for c, e, s, tag in self.boundary:
            kind, given = self.bcs[tag]
            given = np.asarray(given, dtype=np.float64)
            xi, _, wf, normal = self._face_points(c, e, s)
            phi = self._basis(xi)
            uc = phi @ u[c]
            for p in range(len(xi)):
                n = normal[p]
                if kind == WALL:
                    fl = np.zeros(5)
                    fl[1:4] = pressure(uc[p]) * n
                elif kind == OUTLET:
                    fl = euler_flux(uc[p]) @ n
                elif kind == INLET:
                    fl = euler_flux(given) @ n
                elif kind in (SUB_INLET, SUB_OUTLET):
                    fl = boundary(kind, _frame(n), uc[p], given)
                else:
                    fl = upwind(_frame(n), uc[p], given)
                r[c] -= np.outer(phi[p], fl * wf[p])
        # diagonal (collocated) mass
        r /= wq[:, :, None]
        return r.reshape(-1)
