"""mcfd_face_fluxes_noalloc (the entry point that allocates no device memory)
against mcfd_face_fluxes, the allocating entry point it is a variant of.

Both run the same kernel on the same inputs, so the fluxes and the exact
solver's per-point states must be identical bit for bit.  Each solver / form /
kind case makes two calls in a row, the second with fewer points, so the
second call runs on a staging buffer that still holds the first call's data;
one call takes exactly MCFD_NOALLOC_MAX_POINTS points, and one more point is
rejected with MCFD_EINVAL.
"""
import ctypes

import numpy as np
import pytest

from pvcriemannsolvers_amd import _capi as C

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    frames = np.ascontiguousarray(q * np.sign(np.linalg.det(q))[:, None, None]).reshape(n, 9)

    def state():
        rho, p = rng.uniform(0.1, 2, n), rng.uniform(0.05, 5, n)
        vel = rng.uniform(-2, 2, (n, 3))
        return np.ascontiguousarray(np.column_stack([rho, rho[:, None] * vel,
                                                     p / (GAMMA - 1) + 0.5 * rho * (vel ** 2).sum(1)]))
    return frames, state(), state()


def call(fn, riemann, form, kind, frames, left, right, star):
    n = len(left)
    out = np.zeros((n, 5))
    rc = fn(riemann, form, kind, n, C.ptr(frames, C._dp), C.ptr(left, C._dp),
            C.ptr(right, C._dp) if right is not None else None,
            star.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if star is not None else None, C.ptr(out, C._dp))
    C.check(rc)
    return out


CASES = [(C.RIEMANN_EXACT, C.FLUX_FORM_FR, C.FLUX_UPWIND), (C.RIEMANN_HLLC, C.FLUX_FORM_DG, C.FLUX_UPWIND),
         (C.RIEMANN_ROE, C.FLUX_FORM_FR, C.FLUX_UPWIND), (C.RIEMANN_ROE, C.FLUX_FORM_DG, C.FLUX_UPWIND),
         (C.RIEMANN_HLLC, C.FLUX_FORM_FR, C.FLUX_INVISCID_WALL),
         (C.RIEMANN_HLLC, C.FLUX_FORM_FR, C.FLUX_SUPERSONIC_OUTLET)]


@pytest.mark.parametrize("riemann,form,kind", CASES)
def test_noalloc_matches_allocating_entry_point(riemann, form, kind):
    lib = C.load()
    for call_id, n in enumerate((5000, 1200)):  # the second call is smaller
        frames, left, right = inputs(n, 10 * riemann + call_id)
        r = right if kind == C.FLUX_UPWIND else None
        star_a = (np.arange(n) % 2).astype(np.uint8)
        star_b = star_a.copy()
        want = call(lib.mcfd_face_fluxes, riemann, form, kind, frames, left, r, star_a)
        got = call(lib.mcfd_face_fluxes_noalloc, riemann, form, kind, frames, left, r, star_b)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(star_b, star_a)


def test_noalloc_limit():
    lib = C.load()
    n = C.NOALLOC_MAX_POINTS
    frames, left, right = inputs(n + 1, 3)
    # exactly the maximum: fresh solvers (no state array)
    want = call(lib.mcfd_face_fluxes, C.RIEMANN_HLLC, C.FLUX_FORM_FR, C.FLUX_UPWIND, frames[:n], left[:n],
                right[:n].copy(), None)
    got = call(lib.mcfd_face_fluxes_noalloc, C.RIEMANN_HLLC, C.FLUX_FORM_FR, C.FLUX_UPWIND,
               np.ascontiguousarray(frames[:n]), np.ascontiguousarray(left[:n]), np.ascontiguousarray(right[:n]),
               None)
    np.testing.assert_array_equal(got, want)
    # one point more: rejected before anything runs
    out = np.zeros((n + 1, 5))
    rc = lib.mcfd_face_fluxes_noalloc(C.RIEMANN_HLLC, C.FLUX_FORM_FR, C.FLUX_UPWIND, n + 1, C.ptr(frames, C._dp),
                                      C.ptr(left, C._dp), C.ptr(right, C._dp), None, C.ptr(out, C._dp))
    assert rc == C.MCFD_EINVAL
    assert b"MCFD_NOALLOC_MAX_POINTS" in lib.mcfd_last_error()
