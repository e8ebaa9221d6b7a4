"""Restart files: a solution column with the time and the layout it belongs to.

The reference writes the solution coefficients of every cell as CGNS fields
and reads them back to restart (mesh/part.hpp:1070-1208, ``WriteSolutions`` /
``ReadSolutions``); CGNS is not available to this library, so a restart is
a NumPy .npz archive holding the column in the drop-in layout
column[(cell * N + i) * 5 + k], the time, and the scheme, degree, cell
count and unknowns per cell that a reader checks before it sets the column.
"""
from __future__ import annotations

import numpy as np


def write_restart(fe, path: str, t: float) -> None:
    """Saves `fe`'s current solution and the time `t` to `path` (.npz)."""
    np.savez(path, column=fe.get_solution_column(), time=np.float64(t),
             scheme=np.array(fe.scheme), degree=np.int64(fe.degree), n_cells=np.int64(fe.n_cells),
             n_per_cell=np.int64(fe.n_per_cell))


def read_restart(fe, path: str) -> float:
    """Sets `fe`'s solution from a restart file written for the same layout
    and returns its time; raises ValueError if the layout differs.  The
    column is restored as stored: no limiter pass runs on it, as the
    reference's restart reads the coefficients back into the cells
    (demo/euler/sourceless.cpp:167-171)."""
    with np.load(path, allow_pickle=False) as z:
        layout = (str(z["scheme"]), int(z["degree"]), int(z["n_cells"]), int(z["n_per_cell"]))
        want = (fe.scheme, fe.degree, fe.n_cells, fe.n_per_cell)
        if layout != want:
            raise ValueError(f"restart file layout {layout} does not match the solver's {want}")
        fe.restore_solution_column(z["column"])
        return float(z["time"])
