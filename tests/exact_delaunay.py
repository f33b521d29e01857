"""Exact periodic Delaunay one-rings by brute force (test helper).

Points live on the integer grid (coordinates k / 2^10).  For every point v,
all periodic images within distance D are candidates; a counter-clockwise
triangle (v, a, b) is Delaunay iff no candidate lies strictly inside its
circumcircle, decided with the textbook lifted 3x3 in-circle determinant in
exact int64 arithmetic (|terms| < 2^51 for D <= 3 in units of 2^10).  The
ring of v follows the successor map a -> b of its Delaunay triangles.
"""
from __future__ import annotations

import numpy as np

SCALE = 1 << 10


def grid_points(n, lx_cells, ly_cells, seed):
    """n points with coordinates k/2^10 in [0,lx) x [0,ly) (integer lx, ly)."""
    rng = np.random.default_rng(seed)
    X = rng.integers(0, lx_cells * SCALE, size=n)
    Y = rng.integers(0, ly_cells * SCALE, size=n)
    return np.stack([X, Y], axis=1).astype(np.int64)


def _incircle(p, q, r, d):
    """> 0 iff d strictly inside the circle through CCW p, q, r (int64)."""
    ax, ay = p[..., 0] - d[..., 0], p[..., 1] - d[..., 1]
    bx, by = q[..., 0] - d[..., 0], q[..., 1] - d[..., 1]
    cx, cy = r[..., 0] - d[..., 0], r[..., 1] - d[..., 1]
    a2, b2, c2 = ax * ax + ay * ay, bx * bx + by * by, cx * cx + cy * cy
    return (ax * (by * c2 - b2 * cy) - ay * (bx * c2 - b2 * cx) + a2 * (bx * cy - by * cx))


def rings(P, lx_cells, ly_cells, D=3.0):
    n = P.shape[0]
    Lx, Ly = lx_cells * SCALE, ly_cells * SCALE
    R2 = int((D * SCALE) ** 2)
    out = []
    for v in range(n):
        d = P[None, :, :] - P[v][None, None, :]
        shifts = np.array([[sx * Lx, sy * Ly] for sx in (-1, 0, 1) for sy in (-1, 0, 1)])
        cand = (d + shifts[:, None, :]).reshape(-1, 2)
        ids = np.tile(np.arange(n), 9)
        keep = (cand[:, 0] ** 2 + cand[:, 1] ** 2 < R2) & ~((cand[:, 0] == 0) & (cand[:, 1] == 0))
        C, I = cand[keep], ids[keep]
        k = C.shape[0]
        zero = np.zeros(2, dtype=np.int64)
        succ = {}
        for ia in range(k):
            a = C[ia]
            cross = a[0] * C[:, 1] - a[1] * C[:, 0]
            for ib in np.nonzero(cross > 0)[0]:
                b = C[ib]
                inc = _incircle(np.broadcast_to(zero, C.shape), np.broadcast_to(a, C.shape),
                                np.broadcast_to(b, C.shape), C)
                inc[[ia, ib]] = -1
                if np.all(inc < 0):
                    assert ia not in succ, "two Delaunay triangles share a directed edge"
                    succ[ia] = ib
                    # circle must lie inside the candidate disk (completeness)
                    ax, ay, bx, by = (float(a[0]), float(a[1]), float(b[0]), float(b[1]))
                    den = 2.0 * (ax * by - ay * bx)
                    qx = ((ax * ax + ay * ay) * by - (bx * bx + by * by) * ay) / den
                    qy = ((bx * bx + by * by) * ax - (ax * ax + ay * ay) * bx) / den
                    assert 4.0 * (qx * qx + qy * qy) < R2, "increase D"
        start = next(iter(succ))
        ring, cur = [], start
        while True:
            ring.append(int(I[cur]))
            cur = succ[cur]
            if cur == start:
                break
        assert len(ring) == len(succ)
        out.append(ring)
    return out
