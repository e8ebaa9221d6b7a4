"""MI355X-native high-order residual and SSP-RK time stepping (miniCFD hot path).

Host surface mirroring the reference's namespaces:
  * :mod:`.temporal`  -- mini::temporal (System, Euler, RungeKutta)
  * :mod:`.spatial`   -- mini::spatial (FiniteElement on libminicfd)
  * :mod:`.meshgen`   -- synthetic seeded meshes in the reference's input form
The HIP library is loaded on first use by :mod:`._capi`.
"""
from . import meshgen, spatial, temporal  # noqa: F401

__all__ = ["meshgen", "spatial", "temporal"]
