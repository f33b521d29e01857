"""Audit of the FP32 filter's certified bound, decision by decision.

The host build of the device routines with GDT_AUDIT_FILTER (tests/cpu_emul/
emul.cpp) records every decision the FP32 filter takes from its FP32 value
alone -- "c is left of v->a", "c is inside circle (v, a, b)", "c is closer
than b", and their negations -- together with the FP64 displacements the
FP64 path evaluates the same predicate on.  Each claimed sign is checked here
against the EXACT sign of that predicate (rational arithmetic on the FP64
values; numpy FP64 with a rigorous error bound first, Fraction where that
bound cannot decide).  A claim whose exact sign differs, or is zero, is a bound
that is too tight: the filtered path could then take a decision the FP64 path
does not.

Inputs are adversarial for the filter: square and triangular lattices (four
or more points on one circle), jittered by 0 (exact FP64 ties) and by 1e-13
to 1e-6 per coordinate, and clusters of eight points on small circles; both
the LDS tile routine and the per-point routine, FP32-filtered.
"""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from gpudelaunaytriangulation_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_emul", "emul.cpp")
CSRC = os.path.join(ROOT, "gpudelaunaytriangulation_amd", "csrc")
REC = np.dtype([("kind", "<i4"), ("sign", "<i4"), ("ax", "<f8"), ("ay", "<f8"), ("bx", "<f8"),
                ("by", "<f8"), ("cx", "<f8"), ("cy", "<f8")])
ORIENT, INCIRCLE, DIST = 0, 1, 2
U = 2.0 ** -53


def build_audit(path, extra=()):
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-DGDT_AUDIT_FILTER",
                    *extra, f"-I{CSRC}", SRC, "-o", path], check=True)
    return path


@pytest.fixture(scope="module")
def audit_exe(tmp_path_factory):
    return build_audit(str(tmp_path_factory.mktemp("audit") / "emul_audit"))


def square_lattice(m, eps, seed):
    j, i = np.divmod(np.arange(m * m), m)
    xy = np.stack([i + 0.5, j + 0.5], 1).astype(np.float64)
    if eps:
        xy += eps * inputs.gaussian(seed, 2 * m * m).reshape(-1, 2)
    return inputs.wrap_into_box(xy, float(m), float(m)), float(m), float(m)


def rotated_lattice(eps, seed):
    """The unit square lattice rotated by atan(4/3) and shifted by an offset
    that FP32 and FP64 cannot represent: periodic in a 35 x 35 box (1,225
    points), every lattice square cocircular, so the FP32 values of its
    in-circle predicates are rounding noise."""
    i, j = np.meshgrid(np.arange(-60, 61), np.arange(-60, 61))
    i, j = i.ravel(), j.ravel()
    u5, v5 = 3 * i - 4 * j, 4 * i + 3 * j          # 5x, 5y: exact integers
    keep = (u5 >= 0) & (u5 < 175) & (v5 >= 0) & (v5 < 175)
    xy = np.stack([u5[keep] / 5.0 + 0.123456789, v5[keep] / 5.0 + 0.0987654321], 1)
    assert xy.shape[0] == 1225
    if eps:
        xy = xy + eps * inputs.gaussian(seed, xy.size).reshape(-1, 2)
    return inputs.wrap_into_box(xy, 35.0, 35.0), 35.0, 35.0


def tri_lattice(m, eps, seed):
    lx, ly = inputs.lattice_box(m, m)
    return inputs.lattice_points(m, m, eps, seed=seed), lx, ly


def circle_clusters(eps, seed):
    """Eight points on each circle of radius 1 around the nodes of an 11x11
    grid of spacing 3 (box 33): cocircular octets, jittered by eps."""
    g = np.arange(11) * 3.0 + 1.5
    cx, cy = np.meshgrid(g, g)
    t = np.arange(8) * (np.pi / 4)
    xy = np.stack([(cx.ravel()[:, None] + np.cos(t)[None, :]).ravel(),
                   (cy.ravel()[:, None] + np.sin(t)[None, :]).ravel()], 1)
    if eps:
        xy = xy + eps * inputs.gaussian(seed, xy.size).reshape(-1, 2)
    return inputs.wrap_into_box(xy, 33.0, 33.0), 33.0, 33.0


def run_audit(exe, tmp, xy, lx, ly, tiled, mode):
    n = xy.shape[0]
    pin, pout, paud = (os.path.join(tmp, f) for f in ("p.bin", "o.bin", "a.bin"))
    np.ascontiguousarray(xy, dtype=np.float64).tofile(pin)
    subprocess.run([exe, pin, str(n), repr(lx), repr(ly), "1.0", str(tiled), "32", "32", pout,
                    str(mode), paud], check=True, capture_output=True)
    return np.fromfile(paud, dtype=REC)


def exact_value(kind, r):
    F = [Fraction(float(r[k])) for k in ("ax", "ay", "bx", "by", "cx", "cy")]
    ax, ay, bx, by, cx, cy = F
    if kind == ORIENT:
        return ax * cy - ay * cx
    cc, bb = cx * cx + cy * cy, bx * bx + by * by
    if kind == DIST:
        return cc - bb
    aa = ax * ax + ay * ay
    k0, kx, ky = ax * by - ay * bx, bb * ay - aa * by, aa * bx - bb * ax
    return k0 * cc + kx * cx + ky * cy


def check_claims(rec):
    """Indices of claims whose exact sign differs from the claim (or is 0)."""
    ax, ay, bx, by, cx, cy = (rec[k] for k in ("ax", "ay", "bx", "by", "cx", "cy"))
    kind, sign = rec["kind"], rec["sign"]
    val = np.zeros(len(rec))
    bound = np.full(len(rec), np.inf)
    o = kind == ORIENT
    val[o] = ax[o] * cy[o] - ay[o] * cx[o]
    bound[o] = 4 * U * (np.abs(ax[o] * cy[o]) + np.abs(ay[o] * cx[o]))
    d = kind == DIST
    cc = cx * cx + cy * cy
    bb = bx * bx + by * by
    val[d] = cc[d] - bb[d]
    bound[d] = 8 * U * (cc[d] + bb[d])
    q = kind == INCIRCLE
    aa = ax * ax + ay * ay
    k0 = ax * by - ay * bx
    kx = bb * ay - aa * by
    ky = aa * bx - bb * ax
    val[q] = (k0 * cc + kx * cx + ky * cy)[q]
    perm = ((np.abs(ax * by) + np.abs(ay * bx)) * cc + (bb * np.abs(ay) + aa * np.abs(by)) *
            np.abs(cx) + (aa * np.abs(bx) + bb * np.abs(ax)) * np.abs(cy))
    bound[q] = 16 * U * perm[q]
    sure = np.abs(val) > bound
    bad = list(np.flatnonzero(sure & (np.sign(val) != sign)))
    for i in np.flatnonzero(~sure):
        e = exact_value(int(kind[i]), rec[i])
        if e == 0 or (e > 0) != (sign[i] > 0):
            bad.append(int(i))
    return bad, int((~sure).sum())


CASES = [("square", e) for e in (0.0, 1e-13, 1e-11, 1e-9, 1e-6)] + \
        [("rotated", e) for e in (0.0, 1e-13, 1e-10, 1e-7)] + \
        [("triangular", e) for e in (0.0, 1e-13, 1e-10, 1e-7)] + \
        [("circles", e) for e in (0.0, 1e-12, 1e-8)]


@pytest.mark.parametrize("shape,eps", CASES)
@pytest.mark.parametrize("routine", ["tile", "point"])
def test_filter_claims_match_exact_signs(audit_exe, tmp_path, shape, eps, routine):
    if shape == "square":
        xy, lx, ly = square_lattice(32, eps, seed=71)
    elif shape == "rotated":
        xy, lx, ly = rotated_lattice(eps, seed=74)
    elif shape == "triangular":
        xy, lx, ly = tri_lattice(32, eps, seed=72)
    else:
        xy, lx, ly = circle_clusters(eps, seed=73)
    tiled, mode = (1, 1) if routine == "tile" else (0, 0)
    rec = run_audit(audit_exe, str(tmp_path), xy, lx, ly, tiled, mode)
    kinds = np.bincount(rec["kind"], minlength=3)
    assert len(rec) > 1000
    if eps > 0:     # (exact ties defer every point at its first step: distance claims only)
        assert kinds[ORIENT] > 1000 and kinds[INCIRCLE] > 100 and kinds[DIST] > 100, kinds
    bad, exact_checked = check_claims(rec)
    assert not bad, (f"{len(bad)} of {len(rec)} FP32 claims contradict the exact sign, "
                     f"first: {rec[bad[0]]}")


def test_audit_catches_a_wrong_claim():
    """The check is sharp: a claim with its sign flipped, and a claim on an
    exactly zero predicate (cocircular integer points), are both caught."""
    rec = np.zeros(3, dtype=REC)
    rec[0] = (INCIRCLE, -1, 1.0, 0.0, 0.0, 1.0, 0.5, 0.5)    # c strictly inside: claim ok
    rec[1] = (INCIRCLE, 1, 1.0, 0.0, 0.0, 1.0, 0.5, 0.5)     # flipped
    rec[2] = (INCIRCLE, -1, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0)    # (1,1) on the circle: exact 0
    bad, _ = check_claims(rec)
    assert bad == [1, 2]


@pytest.mark.parametrize("routine", ["tile", "point"])
def test_audit_catches_a_too_tight_bound(tmp_path, routine):
    """Built with the FP32 unit roundoff taken 1000x too small, the filter's
    bounds are too tight, and on the rotated lattice (FP32 in-circle values of
    its cocircular squares are rounding noise) the audit finds claims that
    contradict the exact sign."""
    exe = build_audit(str(tmp_path / "emul_tight"), ["-DGDT_U32=5.9604644775390625e-11f"])
    xy, lx, ly = rotated_lattice(0.0, seed=74)
    tiled, mode = (1, 1) if routine == "tile" else (0, 0)
    rec = run_audit(exe, str(tmp_path), xy, lx, ly, tiled, mode)
    bad, _ = check_claims(rec)
    assert len(bad) > 10
