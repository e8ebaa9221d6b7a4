"""ctypes binding of libminicfd (include/minicfd.h).

The library is built in-tree by ``pvcriemannsolvers_amd/csrc/Makefile`` into
``pvcriemannsolvers_amd/_lib/libminicfd.so``.  There is no fallback: if the
library is missing, loading fails with an error naming the build step.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libminicfd.so")

MCFD_OK, MCFD_EINVAL, MCFD_EHIP, MCFD_ESTATE, MCFD_ENOTSUP = 0, 1, 2, 3, 4
HEXA_8, TETRA_4 = 0, 1
SCHEME_FR_GENERAL, SCHEME_DG_GENERAL = 0, 1
NODES_LEGENDRE, NODES_LOBATTO = 0, 1
RIEMANN_EXACT, RIEMANN_HLLC, RIEMANN_ROE = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
FLAG_HOST_ONLY = 1
LIMITER_NONE, LIMITER_WENO_FLAGGED, LIMITER_WENO_ALL, LIMITER_WENO_EIGEN, LIMITER_WENO_EIGEN_ALL = 0, 1, 2, 3, 4
BC_SMART, BC_INVISCID_WALL, BC_SUPERSONIC_OUTLET, BC_SUPERSONIC_INLET = 0, 1, 2, 3
BC_SUBSONIC_INLET, BC_SUBSONIC_OUTLET = 4, 5
FLUX_UPWIND, FLUX_INVISCID_WALL, FLUX_SUPERSONIC_OUTLET, FLUX_SUBSONIC_INLET, FLUX_SUBSONIC_OUTLET = 0, 1, 2, 3, 4
FLUX_FORM_FR, FLUX_FORM_DG = 0, 1
NOALLOC_MAX_POINTS = 16384
STARTUP_AVERAGE, STARTUP_MAJORITY = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


class MeshDesc(ctypes.Structure):
    _fields_ = [("cell_type", ctypes.c_int32), ("n_nodes", ctypes.c_int64), ("coords", _dp),
                ("n_cells", ctypes.c_int64), ("cell_nodes", _i64p), ("n_bfaces", ctypes.c_int64),
                ("bface_nodes", _i64p), ("bface_tags", _i32p), ("period", ctypes.c_double * 3)]


class SchemeDesc(ctypes.Structure):
    _fields_ = [("scheme", ctypes.c_int32), ("degree", ctypes.c_int32), ("nodes", ctypes.c_int32),
                ("riemann", ctypes.c_int32), ("precision", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("c_next", ctypes.c_double), ("limiter", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("weno_w0", ctypes.c_double), ("weno_eps", ctypes.c_double)]


class BcDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("state", ctypes.c_double * 5)]


# mcfd_bc_function: states[n][5] = f(xyz[n][3], t)
BC_FUNCTION = ctypes.CFUNCTYPE(None, ctypes.c_int64, _dp, ctypes.c_double, _dp, ctypes.c_void_p)

# Every function declared in include/minicfd.h with its ctypes signature.
SIGNATURES = {
    "mcfd_create": [ctypes.POINTER(MeshDesc), ctypes.POINTER(SchemeDesc), ctypes.POINTER(BcDesc),
                    ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)],
    "mcfd_destroy": [ctypes.c_void_p],
    "mcfd_last_error": [],
    "mcfd_sizes": [ctypes.c_void_p, _i64p, _i32p, _i64p],
    "mcfd_set_time": [ctypes.c_void_p, ctypes.c_double],
    "mcfd_set_solution_column": [ctypes.c_void_p, _dp],
    "mcfd_get_solution_column": [ctypes.c_void_p, _dp],
    "mcfd_restore_solution_column": [ctypes.c_void_p, _dp],
    "mcfd_get_residual_column": [ctypes.c_void_p, _dp],
    "mcfd_rk3_update": [ctypes.c_void_p, ctypes.c_double, ctypes.c_double],
    "mcfd_rk3_stage": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double],
    "mcfd_get_time_step": [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp],
    "mcfd_sample_points": [ctypes.c_void_p, _i32p, _dp],
    "mcfd_approximate": [ctypes.c_void_p, _dp],
    "mcfd_evaluate": [ctypes.c_void_p, _dp, ctypes.c_int32, _dp, _dp, _dp],
    "mcfd_face_count": [ctypes.c_void_p, _i64p, _i64p],
    "mcfd_faces": [ctypes.c_void_p, _i64p, _i64p, _i32p, _i32p, _i32p],
    "mcfd_limiter_flags": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), _i64p],
    "mcfd_limiter_reconstruct": [ctypes.c_void_p],
    "mcfd_startup_limiter": [ctypes.c_void_p, ctypes.c_int32],
    "mcfd_boundary_points": [ctypes.c_void_p, ctypes.c_int32, _i64p, _dp],
    "mcfd_set_boundary_function": [ctypes.c_void_p, ctypes.c_int32, BC_FUNCTION, ctypes.c_void_p],
    "mcfd_set_boundary_values": [ctypes.c_void_p, ctypes.c_int32, _dp],
    "mcfd_face_fluxes": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _dp, _dp, _dp,
                         ctypes.POINTER(ctypes.c_uint8), _dp],
    "mcfd_face_fluxes_noalloc": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _dp, _dp, _dp,
                                 ctypes.POINTER(ctypes.c_uint8), _dp],
    "mcfd_synchronize": [ctypes.c_void_p],
    "mcfd_time_rk3": [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, _dp, _dp, _dp],
    "mcfd_batch_create": [ctypes.c_int32, ctypes.POINTER(MeshDesc), _i32p, ctypes.POINTER(SchemeDesc),
                          ctypes.POINTER(BcDesc), _i32p, ctypes.POINTER(ctypes.c_void_p)],
    "mcfd_batch_destroy": [ctypes.c_void_p],
    "mcfd_batch_case_sizes": [ctypes.c_void_p, ctypes.c_int32, _i64p, _i32p],
    "mcfd_batch_sample_points": [ctypes.c_void_p, ctypes.c_int32, _dp],
    "mcfd_batch_set_solution_column": [ctypes.c_void_p, ctypes.c_int32, _dp],
    "mcfd_batch_get_solution_column": [ctypes.c_void_p, ctypes.c_int32, _dp],
    "mcfd_batch_get_time_steps": [ctypes.c_void_p, ctypes.c_double, _dp],
    "mcfd_batch_set_time_steps": [ctypes.c_void_p, _dp],
    "mcfd_batch_autotune": [ctypes.c_void_p, ctypes.c_int32, _i32p, _i32p],
    "mcfd_batch_set_mode": [ctypes.c_void_p, ctypes.c_int32, _i32p],
    "mcfd_batch_rk3_steps": [ctypes.c_void_p, ctypes.c_int32],
    "mcfd_batch_synchronize": [ctypes.c_void_p],
    "mcfd_batch_time_rk3": [ctypes.c_void_p, ctypes.c_int32, _dp],
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C pvcriemannsolvers_amd/csrc` "
                           "or __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "mcfd_last_error" else ctypes.c_int
    _lib = lib
    return lib


class MinicfdError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"minicfd error {code}: {message}")
        self.code = code


def check(rc: int):
    if rc != MCFD_OK:
        raise MinicfdError(rc, load().mcfd_last_error().decode())


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)
