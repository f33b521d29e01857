"""Exact Delaunay check of rings for arbitrary FP64 points (test helper).

Every double is an integer multiple of a power of two, so the points and the
box sides scale exactly to Python integers by one common power of two; the
periodic displacements, orientations and in-circle determinants are then
evaluated in exact integer arithmetic, whatever the inputs' rounding.  A
triangle (v, a, b) of the rings must be counter-clockwise and have no point
strictly inside its circumcircle; a point exactly on it is decided by the
symbolic perturbation the library uses for exact ties (lifted coordinate
|p|^2 + eps^(1 + caller index), the term of the smallest caller index
dominating).  Zero violations with consistent rings pin the triangulation of
the perturbed points, which is unique.
"""
from __future__ import annotations

from fractions import Fraction


def to_integers(xy, lx, ly):
    """Points and box sides as integers in units of 2^-k (exact)."""
    vals = [Fraction(float(v)) for v in list(xy.reshape(-1)) + [lx, ly]]
    den = 1
    for f in vals:
        den = max(den, f.denominator)          # every denominator is a power of two
    ints = [int(f * den) for f in vals]
    n = xy.shape[0]
    P = [(ints[2 * i], ints[2 * i + 1]) for i in range(n)]
    return P, ints[-2], ints[-1], den


def _round_div(a, b):
    """round(a / b) for integers (ties away from zero do not occur here)."""
    q, r = divmod(2 * a + b, 2 * b)
    return q


def _incircle(p, q, r):
    """Rows p, q, r (the CCW triangle's vertices relative to the test point):
    > 0 iff the test point is strictly inside their circle."""
    ax, ay = p
    bx, by = q
    cx, cy = r
    a2, b2, c2 = ax * ax + ay * ay, bx * bx + by * by, cx * cx + cy * cy
    return (ax * (by * c2 - b2 * cy) - ay * (bx * c2 - b2 * cx) + a2 * (bx * cy - by * cx))


def _sos(coef, ids):
    for _, c in sorted(zip(ids, coef)):
        if c != 0:
            return 1 if c > 0 else -1
    return 0


def exact_violations(xy, lx, ly, nbrs, counts, D):
    """Triangles of the rings that are not Delaunay (exactly, ties perturbed).
    D: distance (box units) within which every point is tested against the
    circles of v's triangles; the box must exceed 2 D + the circles' size."""
    P, Lx, Ly, den = to_integers(xy, lx, ly)
    Dm = int(D * den)
    n = len(P)
    bad = 0
    for v in range(n):
        pv = P[v]
        disp = {}
        near = []
        for u in range(n):
            if u == v:
                continue
            dx, dy = P[u][0] - pv[0], P[u][1] - pv[1]
            dx -= Lx * _round_div(dx, Lx)
            dy -= Ly * _round_div(dy, Ly)
            disp[u] = (dx, dy)
            if abs(dx) <= Dm and abs(dy) <= Dm:
                near.append(u)
        c = int(counts[v])
        for m in range(c):
            a, b = int(nbrs[v][m]), int(nbrs[v][(m + 1) % c])
            A, B = disp[a], disp[b]
            if A[0] * B[1] - A[1] * B[0] <= 0:
                bad += 1
                continue
            for t in near:
                if t == a or t == b:
                    continue
                K = disp[t]
                p = (-K[0], -K[1])
                q = (A[0] - K[0], A[1] - K[1])
                r = (B[0] - K[0], B[1] - K[1])
                d0 = _incircle(p, q, r)
                if d0 > 0:
                    bad += 1
                elif d0 == 0:
                    # determinant of rows v-t, a-t, b-t with lifted column; the
                    # perturbation adds (e_p - e_t) times each cofactor
                    cp = q[0] * r[1] - q[1] * r[0]
                    cq = -(p[0] * r[1] - p[1] * r[0])
                    cr = p[0] * q[1] - p[1] * q[0]
                    if _sos([cp, cq, cr, -(cp + cq + cr)], [v, a, b, t]) >= 0:
                        bad += 1
    return bad
