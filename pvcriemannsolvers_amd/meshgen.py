"""Synthetic seeded meshes in the reference's input form.

The reference reads CGNS meshes partitioned by METIS (mesh/part.hpp:453-476).
Here the meshes are generated in memory and handed over as plain arrays:

* ``coords``      float64 [n_nodes, 3]
* ``cells``       int64   [n_cells, npe], CGNS node order (HEXA_8 / TETRA_4)
* ``bface_nodes`` int64   [n_bfaces, face_npe]
* ``bface_tags``  int32   [n_bfaces]; a boundary tag indexes the BC table,
  a negative tag marks a periodic face to be paired across the box.
* ``period``      float64 [3]; box length of each periodic direction, else 0.

Boundary tags of a box: 0 = x-, 1 = x+, 2 = y-, 3 = y+, 4 = z-, 5 = z+.
"""
from __future__ import annotations

import numpy as np


def box_hex(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
            periodic=(False, False, False), shuffle_seed: int | None = None, jitter: float = 0.0,
            seed: int = 0) -> dict:
    """Structured hexahedral box; cell order optionally shuffled by seed.

    ``jitter`` moves every vertex off the box boundary by U(-jitter*h,
    +jitter*h) per coordinate (h the spacing in that direction), seeded, so
    the cells become general trilinear hexahedra with non-planar faces while
    the box sides stay planar (and periodic partners stay translates).
    """
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    xs = [np.linspace(lo[d], hi[d], n + 1) for d, n in enumerate((nx, ny, nz))]
    # node id = (k * (ny + 1) + j) * (nx + 1) + i
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([xs[0][ii], xs[1][jj], xs[2][kk]], axis=-1).reshape(-1, 3)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        h = (hi - lo) / np.array([nx, ny, nz])
        inner = np.all((coords > lo + 0.5 * h) & (coords < hi - 0.5 * h), axis=1)
        coords[inner] += rng.uniform(-jitter, jitter, (int(inner.sum()), 3)) * h

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    cells = np.stack([
        nid(ci, cj, ck), nid(ci + 1, cj, ck), nid(ci + 1, cj + 1, ck), nid(ci, cj + 1, ck),
        nid(ci, cj, ck + 1), nid(ci + 1, cj, ck + 1), nid(ci + 1, cj + 1, ck + 1), nid(ci, cj + 1, ck + 1),
    ], axis=1).astype(np.int64)
    if shuffle_seed is not None:
        perm = np.random.default_rng(shuffle_seed).permutation(len(cells))
        cells = cells[perm]

    faces, tags = [], []
    a, b = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    a, b = a.ravel(), b.ravel()
    for side, i in ((0, 0), (1, nx)):
        faces.append(np.stack([nid(i, a, b), nid(i, a + 1, b), nid(i, a + 1, b + 1), nid(i, a, b + 1)], 1))
        tags.append(np.full(len(a), -1 if periodic[0] else side))
    a, b = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    a, b = a.ravel(), b.ravel()
    for side, j in ((2, 0), (3, ny)):
        faces.append(np.stack([nid(a, j, b), nid(a + 1, j, b), nid(a + 1, j, b + 1), nid(a, j, b + 1)], 1))
        tags.append(np.full(len(a), -1 if periodic[1] else side))
    a, b = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    a, b = a.ravel(), b.ravel()
    for side, k in ((4, 0), (5, nz)):
        faces.append(np.stack([nid(a, b, k), nid(a + 1, b, k), nid(a + 1, b + 1, k), nid(a, b + 1, k)], 1))
        tags.append(np.full(len(a), -1 if periodic[2] else side))
    period = np.array([hi[d] - lo[d] if periodic[d] else 0.0 for d in range(3)])
    return {
        "cell_type": 0,
        "coords": np.ascontiguousarray(coords, dtype=np.float64),
        "cells": np.ascontiguousarray(cells, dtype=np.int64),
        "bface_nodes": np.ascontiguousarray(np.concatenate(faces), dtype=np.int64),
        "bface_tags": np.ascontiguousarray(np.concatenate(tags), dtype=np.int32),
        "period": period,
    }


# Six tetrahedra per hexahedron around the diagonal from corner 0 to corner 6
# (CGNS HEXA_8 corner numbering); translation-invariant, hence conforming.
_KUHN = ((0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6))


def box_tet(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter: float = 0.0,
            seed: int = 0) -> dict:
    """Hexahedral box split into 6 tetrahedra per hexahedron (TETRA_4, CGNS order).

    ``jitter`` moves every vertex off the boundary by U(-jitter*h, +jitter*h)
    per coordinate (h the spacing in that direction), seeded.  Every tetrahedron
    is right-handed in CGNS order: (x1-x0) x (x2-x0) . (x3-x0) > 0.  Boundary
    triangles carry the box-side tags 0..5 (x-, x+, y-, y+, z-, z+).
    """
    hexa = box_hex(nx, ny, nz, lo, hi)
    coords = hexa["coords"].copy()
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        h = (hi - lo) / np.array([nx, ny, nz])
        inner = np.all((coords > lo + 0.5 * h) & (coords < hi - 0.5 * h), axis=1)
        coords[inner] += rng.uniform(-jitter, jitter, (int(inner.sum()), 3)) * h
    cells = hexa["cells"][:, np.array(_KUHN)].reshape(-1, 4)
    x = coords[cells]
    vol = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
    flip = vol < 0
    cells[flip] = cells[flip][:, [0, 2, 1, 3]]
    # boundary triangles: faces used by one tetrahedron only
    faces = np.concatenate([cells[:, [0, 1, 2]], cells[:, [0, 1, 3]], cells[:, [0, 2, 3]], cells[:, [1, 2, 3]]])
    key = np.sort(faces, axis=1)
    _, first, counts = np.unique(key, axis=0, return_index=True, return_counts=True)
    bfaces = faces[np.sort(first[counts == 1])]
    fx = coords[bfaces]
    tags = np.full(len(bfaces), -1, dtype=np.int32)
    eps = 1e-12 * (1 + np.abs(hi - lo).max())
    for d in range(3):
        tags[np.all(np.abs(fx[:, :, d] - lo[d]) < eps, axis=1)] = 2 * d
        tags[np.all(np.abs(fx[:, :, d] - hi[d]) < eps, axis=1)] = 2 * d + 1
    assert np.all(tags >= 0)
    return {
        "cell_type": 1,
        "coords": np.ascontiguousarray(coords),
        "cells": np.ascontiguousarray(cells, dtype=np.int64),
        "bface_nodes": np.ascontiguousarray(bfaces, dtype=np.int64),
        "bface_tags": tags,
        "period": np.zeros(3),
    }
