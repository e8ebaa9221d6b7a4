"""mini::riemann surface over the device solvers of libminicfd.

``Rotated`` mirrors ``rotated::Euler<Unrotated>`` (riemann/rotated/euler.hpp)
for a set of face points, one solver object per point as the reference's
``FiniteElement`` keeps them (spatial/fem.hpp:95-115): the upwind flux
(:153-159) and the boundary fluxes of an inviscid wall (:160-174), a
supersonic outlet (:144-151) and a subsonic inlet and outlet (:181-280).  Each call runs the stage kernels' own device
code (``mcfd_face_fluxes``); the exact solver's per-point state (the sign of
the last star velocity, riemann/euler/exact.hpp:29) persists across calls.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C

_KIND = {"exact": C.RIEMANN_EXACT, "hllc": C.RIEMANN_HLLC, "roe": C.RIEMANN_ROE}
_FORM = {"fr": C.FLUX_FORM_FR, "dg": C.FLUX_FORM_DG}


class Rotated:
    """``riemann``: "exact", "hllc" or "roe"; ``form``: "fr" (the FR stage
    kernels' evaluation) or "dg" (the modal-DG face pass's)."""

    def __init__(self, riemann: str, n_points: int, form: str = "fr"):
        self.riemann = _KIND[riemann]
        self.form = _FORM[form]
        self.exact_star = np.zeros(int(n_points), np.uint8)

    def _call(self, kind, frames, left, right):
        frames = np.ascontiguousarray(frames, np.float64).reshape(-1, 9)
        left = np.ascontiguousarray(left, np.float64).reshape(-1, 5)
        n = len(left)
        assert len(frames) == n == len(self.exact_star)
        r = None if right is None else np.ascontiguousarray(right, np.float64).reshape(-1, 5)
        out = np.empty((n, 5))
        star = self.exact_star.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        C.check(C.load().mcfd_face_fluxes(self.riemann, self.form, kind, n, C.ptr(frames, C._dp),
                                          C.ptr(left, C._dp), C.ptr(r, C._dp) if r is not None else None,
                                          star, C.ptr(out, C._dp)))
        return out

    def get_flux_upwind(self, frames, left, right):
        """Global flux of the conservative pair (left, right) at each point; the
        first frame row is the normal, from left to right."""
        return self._call(C.FLUX_UPWIND, frames, left, right)

    def get_flux_on_inviscid_wall(self, frames, state):
        return self._call(C.FLUX_INVISCID_WALL, frames, state, None)

    def get_flux_on_supersonic_outlet(self, frames, state):
        return self._call(C.FLUX_SUPERSONIC_OUTLET, frames, state, None)

    def get_flux_on_subsonic_inlet(self, frames, state, given):
        """given = (p_total, u_cos, v_cos, w_cos, T_total) per point (rotated/euler.hpp:181-246)."""
        return self._call(C.FLUX_SUBSONIC_INLET, frames, state, given)

    def get_flux_on_subsonic_outlet(self, frames, state, given):
        """given[4] = the back pressure per point (rotated/euler.hpp:263-280)."""
        return self._call(C.FLUX_SUBSONIC_OUTLET, frames, state, given)
