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
    n = xy.shape[0]
    tiles, ids, shifts = [], [], []
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            p = xy + np.array([sx * lx, sy * ly])
            keep = ((p[:, 0] > -margin) & (p[:, 0] < lx + margin)
                    & (p[:, 1] > -margin) & (p[:, 1] < ly + margin))
            tiles.append(p[keep])
            ids.append(np.nonzero(keep)[0])
    P = np.concatenate(tiles)
    I = np.concatenate(ids)
    tri = Delaunay(P)
    indptr, indices = tri.vertex_neighbor_vertices
    # original points are the (0,0) tile: locate them
    off = 0
    for k, (sx, sy) in enumerate([(a, b) for a in (-1, 0, 1) for b in (-1, 0, 1)]):
        if (sx, sy) == (0, 0):
            base = off
        off += tiles[k].shape[0]
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
