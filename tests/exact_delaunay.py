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


def lattice_points(n, box_cells, seed):
    """n points on the 2^-10 grid inside the parallelogram spanned by the
    integer lattice vectors a = (x11, x21), b = (x12, x22) of
    box_cells = (x11, x12, x21, x22) (int64, units of 2^-10; rejection
    sampling with the exact membership test 0 <= det * u < det per unit-cell
    coordinate, det * u = adj(X) p)."""
    x11, x12, x21, x22 = (int(v) for v in box_cells)
    det = x11 * x22 - x12 * x21
    assert det > 0
    xs = [0, x11, x12, x11 + x12]
    ys = [0, x21, x22, x21 + x22]
    rng = np.random.default_rng(seed)
    out = np.zeros((0, 2), dtype=np.int64)
    while out.shape[0] < n:
        X = rng.integers(min(xs) * SCALE, max(xs) * SCALE, size=4 * n)
        Y = rng.integers(min(ys) * SCALE, max(ys) * SCALE, size=4 * n)
        du = x22 * X - x12 * Y
        dv = -x21 * X + x11 * Y
        m = det * SCALE
        keep = (du >= 0) & (du < m) & (dv >= 0) & (dv < m)
        out = np.concatenate([out, np.stack([X[keep], Y[keep]], 1)])
    return out[:n].astype(np.int64)


def rings(P, lx_cells, ly_cells, D=3.0, box_cells=None, images=1):
    """Exact rings of the points P (integer units of 2^-10) in a periodic
    lx_cells x ly_cells box, or in the general box whose integer lattice
    vectors are the columns of box_cells = (x11, x12, x21, x22); candidate
    images p + i a + j b for |i|, |j| <= images."""
    n = P.shape[0]
    if box_cells is None:
        box_cells = (lx_cells, 0, 0, ly_cells)
    x11, x12, x21, x22 = (int(v) * SCALE for v in box_cells)
    R2 = int((D * SCALE) ** 2)
    out = []
    rng_i = range(-images, images + 1)
    shifts = np.array([[i * x11 + j * x12, i * x21 + j * x22] for i in rng_i for j in rng_i],
                      dtype=np.int64)
    for v in range(n):
        d = P[None, :, :] - P[v][None, None, :]
        cand = (d + shifts[:, None, :]).reshape(-1, 2)
        ids = np.tile(np.arange(n), len(shifts))
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


def _sos_sign(d0, coef, ids):
    """Sign of d0 + sum_j coef[j] eps^(1 + ids[j]) as eps -> 0+ (0 only when
    every term vanishes): the term of the smallest caller index dominates."""
    if d0 != 0:
        return 1 if d0 > 0 else -1
    for j in np.argsort(ids, kind="stable"):
        if coef[j] != 0:
            return 1 if coef[j] > 0 else -1
    return 0


def perturbed_violations(P, lx_cells, ly_cells, nbrs, counts, D=3.0):
    """Triangles of the rings that are not Delaunay for the symbolically
    perturbed points (lifted coordinate |p|^2 + eps^(1 + caller index)).

    Every point v with ring a_0 .. a_{c-1} gives triangles (v, a_m, a_m+1);
    each must be counter-clockwise and have no other point inside or on its
    circle after the perturbation.  The 3x3 in-circle determinant is linear in
    its lifted column: with rows p-d, q-d, r-d (coordinates relative to d) it
    gains (e_p - e_d) C_p + (e_q - e_d) C_q + (e_r - e_d) C_r, C_x the
    cofactors of the lifted column, whose sum is orient(p, q, r).  For perturbed
    points in general position the Delaunay triangulation is unique, so zero
    violations together with consistent rings pin the triangulation.  Box of
    at least 2 D + 1 in units of the spacing (one image of each point within D).
    """
    n = P.shape[0]
    Lx, Ly = int(round(lx_cells * SCALE)), int(round(ly_cells * SCALE))
    Dm = int(D * SCALE)
    bad = 0
    for v in range(n):
        d = P - P[v]
        d[:, 0] -= Lx * np.round(d[:, 0] / Lx).astype(np.int64)
        d[:, 1] -= Ly * np.round(d[:, 1] / Ly).astype(np.int64)
        near = np.nonzero((np.abs(d[:, 0]) <= Dm) & (np.abs(d[:, 1]) <= Dm))[0]
        near = near[near != v]
        C = d[near]
        cv = int(counts[v])
        for m in range(cv):
            a, b = int(nbrs[v, m]), int(nbrs[v, (m + 1) % cv])
            A, B = d[a], d[b]
            if A[0] * B[1] - A[1] * B[0] <= 0:
                bad += 1
                continue
            keep = (near != a) & (near != b)
            K, ik = C[keep], near[keep]
            z = np.zeros_like(K)
            d0 = _incircle(z, np.broadcast_to(A, K.shape), np.broadcast_to(B, K.shape), K)
            inside = d0 > 0
            for t in np.nonzero(d0 == 0)[0]:
                p, q, r = -K[t], A - K[t], B - K[t]
                cp = int(q[0] * r[1] - q[1] * r[0])
                cq = -int(p[0] * r[1] - p[1] * r[0])
                cr = int(p[0] * q[1] - p[1] * q[0])
                s = _sos_sign(0, [cp, cq, cr, -(cp + cq + cr)], [v, a, b, int(ik[t])])
                inside[t] = s >= 0
            bad += int(inside.sum())
    return bad


def ring_inconsistencies(nbrs, counts):
    """Directed ring edges whose triangle is missing from a vertex's ring:
    triangle (v, a, b) counter-clockwise means b is followed by v around a."""
    n = len(counts)
    pos = [dict() for _ in range(n)]
    for i in range(n):
        for m in range(int(counts[i])):
            pos[i][int(nbrs[i, m])] = m
    bad = 0
    for i in range(n):
        ci = int(counts[i])
        for m in range(ci):
            a, b = int(nbrs[i, m]), int(nbrs[i, (m + 1) % ci])
            if b not in pos[a] or i not in pos[a] or (pos[a][b] + 1) % int(counts[a]) != pos[a][i]:
                bad += 1
    return bad


def square_lattice(nx, ny, spacing=1.0):
    """Exact square lattice of the given spacing (a multiple of 2^-9 cells),
    points at the centres of its squares, integer coordinates in units of
    2^-10; periodic box nx * spacing by ny * spacing."""
    h = int(round(spacing * SCALE))
    assert h == spacing * SCALE and h % 2 == 0
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    return np.stack([X.reshape(-1) * h + h // 2, Y.reshape(-1) * h + h // 2], axis=1).astype(np.int64)
