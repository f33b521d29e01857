"""Independent periodic Delaunay rings from scipy's Qhull (test helper).

The points are tiled with their periodic images inside a margin around the
box (the nine-sheet idea of the reference's CGAL wrapper for non-square
boxes, DelaunayCGAL.cpp:148-225), triangulated by Qhull, and each original
point's neighbours are mapped back to original indices and ordered
counter-clockwise by angle.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import Delaunay


def periodic_rings(xy: np.ndarray, lx: float, ly: float, margin: float = 4.0):
    return periodic_rings_box(xy, (lx, 0.0, 0.0, ly), margin)


def periodic_rings_box(xy: np.ndarray, box, margin: float = 4.0):
    """Rings in the general periodic box box = (x11, x12, x21, x22), whose
    lattice vectors are its columns a = (x11, x21), b = (x12, x22): the nine
    sheets p + i a + j b (|i|, |j| <= 1) are kept where their unit-cell
    coordinates lie within `margin` (a distance, at most a box thickness) of
    the unit cell.  Points are first wrapped into the unit cell."""
    x11, x12, x21, x22 = (float(v) for v in box)
    X = np.array([[x11, x12], [x21, x22]])
    Xi = np.linalg.inv(X)
    det = abs(np.linalg.det(X))
    tx = det / np.hypot(x12, x22)       # thickness across the walls along b
    ty = det / np.hypot(x11, x21)
    assert margin < min(tx, ty)
    mx, my = margin / tx, margin / ty
    xy = np.asarray(xy, dtype=np.float64)
    n = xy.shape[0]
    u = xy @ Xi.T
    k = np.floor(u)
    u -= k
    xy = xy - k @ X.T           # (points inside the box are kept as they are)
    tiles, ids = [], []
    sheets = [(a, b) for a in (-1, 0, 1) for b in (-1, 0, 1)]
    for sx, sy in sheets:
        us = u + np.array([sx, sy])
        keep = (us[:, 0] > -mx) & (us[:, 0] < 1 + mx) & (us[:, 1] > -my) & (us[:, 1] < 1 + my)
        tiles.append(xy[keep] + (sx * X[:, 0] + sy * X[:, 1]))
        ids.append(np.nonzero(keep)[0])
    P = np.concatenate(tiles)
    I = np.concatenate(ids)
    tri = Delaunay(P)
    indptr, indices = tri.vertex_neighbor_vertices
    # original points are the (0,0) sheet (all of them: its unit coordinates lie in [0,1))
    base = sum(t.shape[0] for t, sh in zip(tiles, sheets) if sh < (0, 0))
    assert tiles[sheets.index((0, 0))].shape[0] == n
    rings = []
    maxdeg = 0
    for i in range(n):
        v = base + i
        nb = indices[indptr[v]:indptr[v + 1]]
        d = P[nb] - P[v]
        ang = np.arctan2(d[:, 1], d[:, 0])
        order = np.argsort(ang, kind="stable")
        ring = I[nb[order]]
        rings.append(ring)
        maxdeg = max(maxdeg, len(ring))
    counts = np.array([len(r) for r in rings], dtype=np.int32)
    nbrs = np.full((n, maxdeg), -1, dtype=np.int32)
    for i, r in enumerate(rings):
        nbrs[i, :len(r)] = r
    return nbrs, counts
