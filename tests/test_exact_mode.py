"""Exact precision mode (GDT_PREC_EXACT = 128), on the CPU.

* The expansion predicates of gdt_exact.h (tests/cpp/exact_check.cpp, built
  here with g++): exact displacements equal the rational value of
  (x_c - x_v) + m L; the signs of orientation, in-circle and distance
  comparison equal those of exact rational arithmetic (fractions.Fraction) on
  random and adversarial near-degenerate configurations (four points on a
  circle, then moved by a few ulps, some across periodic images), where FP64
  evaluation gets the sign wrong.
* The library's CPU mode in exact precision (the kernels' per-point routine
  gdt_one_ring_exact on the host): on lattices whose spacing is not a dyadic
  number, and on exact lattices moved by a few ulps, its rings are consistent,
  sum to 6N and have no triangle whose circle holds a point, checked in exact
  integer arithmetic with ties perturbed (tests/exact_rational.py); the FP64
  path fails that check on the same inputs.  On generic input the exact mode
  equals the FP64 path (and so the oracle).
* A general box and coordinates far outside the box are refused.
"""
import ctypes
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle
from gpudelaunaytriangulation_amd import PREC_EXACT, PREC_FP64, _lib, inputs, rings

from exact_delaunay import ring_inconsistencies
from exact_rational import exact_violations

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpudelaunaytriangulation_amd", "csrc")


@pytest.fixture(scope="module")
def exact_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("exact") / "exact_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{CSRC}",
                    os.path.join(ROOT, "tests", "cpp", "exact_check.cpp"), "-o", exe], check=True)
    return exe


def _sign(x):
    return int(x > 0) - int(x < 0)


def _exact_disp(v, u, L):
    fx = u - v
    fx -= L * np.rint(fx / L)
    m = np.rint((fx - (u - v)) / L)
    return Fraction(u) - Fraction(v) + int(m) * Fraction(L)


def _expected(row):
    vx, vy, ax, ay, bx, by, cx, cy, lx, ly = row
    A = (_exact_disp(vx, ax, lx), _exact_disp(vy, ay, ly))
    B = (_exact_disp(vx, bx, lx), _exact_disp(vy, by, ly))
    C = (_exact_disp(vx, cx, lx), _exact_disp(vy, cy, ly))
    orient = A[0] * C[1] - A[1] * C[0]
    aa, bb, cc = (A[0] ** 2 + A[1] ** 2, B[0] ** 2 + B[1] ** 2, C[0] ** 2 + C[1] ** 2)
    k0 = A[0] * B[1] - A[1] * B[0]
    kx = bb * A[1] - aa * B[1]
    ky = aa * B[0] - bb * A[0]
    P = k0 * cc + kx * C[0] + ky * C[1]
    return _sign(orient), _sign(P), _sign(cc - bb), A[0]


def _cases(rng, n):
    rows = []
    for t in range(n):
        lx, ly = float(rng.choice([1.0, 3.7, 10.0, 0.3])), float(rng.choice([1.0, 2.9, 10.0, 0.3]))
        cx0, cy0 = rng.uniform(0, lx), rng.uniform(0, ly)
        r = rng.uniform(0.01, 0.2) * min(lx, ly)
        th = np.sort(rng.uniform(0, 2 * np.pi, 4))
        pts = [(cx0 + r * np.cos(a), cy0 + r * np.sin(a)) for a in th]   # nearly cocircular
        if t % 3 == 0:                         # a few ulps off the circle
            pts = [(x + int(rng.integers(-3, 4)) * np.spacing(x), y + int(rng.integers(-3, 4)) * np.spacing(y))
                   for x, y in pts]
        elif t % 3 == 1:                       # generic
            pts = [(x + rng.normal(0, r * 0.1), y + rng.normal(0, r * 0.1)) for x, y in pts]
        # some points given as other periodic images, wrapped into [0, L) or not
        out = []
        for x, y in pts:
            kx, ky = int(rng.integers(-1, 2)), int(rng.integers(-1, 2))
            out += [x + kx * lx, y + ky * ly]
        rows.append(out + [lx, ly])
    return rows


def test_expansion_predicates_match_rationals(exact_check):
    rng = np.random.default_rng(123)
    rows = _cases(rng, 600)
    text = "\n".join(" ".join(float(v).hex() for v in r) for r in rows) + "\n"
    res = subprocess.run([exact_check], input=text, capture_output=True, text=True, check=True)
    lines = res.stdout.strip().split("\n")
    assert len(lines) == len(rows)
    zeros = wrong64 = 0
    for r, line in zip(rows, lines):
        f = line.split()
        so, si, sd, ovf = (int(v) for v in f[:4])
        assert ovf == 0
        eo, ei, ed, ax = _expected(r)
        assert (so, si, sd) == (eo, ei, ed), (r, line)
        assert sum(Fraction(float.fromhex(v)) for v in f[4:]) == ax
        zeros += ei == 0
        # the FP64 value of the in-circle polynomial on FP64 displacements
        vx, vy, axx, ayy, bxx, byy, cxx, cyy, lx, ly = r
        d = lambda v, u, L: (u - v) - L * np.rint((u - v) / L)
        A = (d(vx, axx, lx), d(vy, ayy, ly))
        B = (d(vx, bxx, lx), d(vy, byy, ly))
        C = (d(vx, cxx, lx), d(vy, cyy, ly))
        aa, bb, cc = A[0] ** 2 + A[1] ** 2, B[0] ** 2 + B[1] ** 2, C[0] ** 2 + C[1] ** 2
        P64 = (A[0] * B[1] - A[1] * B[0]) * cc + (bb * A[1] - aa * B[1]) * C[0] + (aa * B[0] - bb * A[0]) * C[1]
        wrong64 += _sign(P64) != ei
    # the near-degenerate cases are where FP64 fails: the test covers them
    assert wrong64 >= 20


def host_tri(xy, lx, ly, prec, cell=1.0, stride=32):
    L = _lib.load()
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    nb = np.full((n, stride), -1, dtype=np.int32)
    cnt = np.zeros(n, dtype=np.int32)
    req = ctypes.c_int(0)
    st = L.gdt_host_triangulate(xy.ctypes.data, n, lx, ly, cell, stride, nb.ctypes.data,
                                cnt.ctypes.data, prec, 4, ctypes.byref(req))
    return st, nb, cnt


def lattice(nx, ny, h, jitter_ulps=0, seed=0):
    """Centres of an nx x ny lattice of spacing h (rounded to FP64), moved by
    up to jitter_ulps ulps per coordinate; periodic box nx h x ny h."""
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    xy = np.stack([(X.reshape(-1) + 0.5) * h, (Y.reshape(-1) + 0.5) * h], 1)
    if jitter_ulps:
        rng = np.random.default_rng(seed)
        k = rng.integers(-jitter_ulps, jitter_ulps + 1, size=xy.shape)
        xy = xy + k * np.spacing(xy)
    return xy, nx * h, ny * h


NEAR = {
    "decimal_lattice_0.1": (lattice(14, 14, 0.1), 0.1),
    "lattice_1.1": (lattice(14, 14, 1.1), 1.1),
    "integer_lattice_3ulp": (lattice(14, 14, 1.0, 3, 1), 1.0),
    "lattice_0.3_1ulp": (lattice(16, 12, 0.3, 1, 2), 0.3),
}


@pytest.mark.parametrize("name", sorted(NEAR))
def test_cpu_exact_mode_is_delaunay(name):
    (xy, lx, ly), cell = NEAR[name]
    n = xy.shape[0]
    st, nb, cnt = host_tri(xy, lx, ly, PREC_EXACT, cell)
    assert st == 0
    assert int(cnt.sum()) == 6 * n
    assert ring_inconsistencies(nb, cnt) == 0
    assert exact_violations(xy, lx, ly, nb, cnt, 3.0 * cell) == 0
    # the FP64 path does not get these right: the check above can fail
    st64, nb64, cnt64 = host_tri(xy, lx, ly, PREC_FP64, cell)
    assert st64 != 0 or exact_violations(xy, lx, ly, nb64, cnt64, 3.0 * cell) > 0


def test_cpu_exact_mode_equals_fp64_on_generic_input():
    for n, lx, ly, seed in ((4096, 64.0, 64.0, 1), (1000, 40.0, 25.0, 7)):
        xy = inputs.uniform_points(n, lx, ly, seed=seed)
        st, nb, cnt = host_tri(xy, lx, ly, PREC_EXACT)
        assert st == 0
        on, oc = oracle.triangulate(xy, lx, ly, 1.0, 32)
        assert rings.ring_hash(nb, cnt) == rings.ring_hash(on, oc)


def test_cpu_exact_mode_refuses_what_it_cannot_bound():
    L = _lib.load()
    xy = inputs.uniform_points(500, 20.0, 20.0, seed=3)
    far = xy.copy()
    far[7, 0] += 100.0                        # five box lengths out
    st, _, _ = host_tri(far, 20.0, 20.0, PREC_EXACT)
    assert st == _lib.GDT_ERR_INVALID
    box = np.array([20.0, 5.0, 0.0, 20.0])   # sheared
    n = xy.shape[0]
    nb = np.full((n, 32), -1, dtype=np.int32)
    cnt = np.zeros(n, dtype=np.int32)
    req = ctypes.c_int(0)
    st = L.gdt_host_triangulate_general(np.ascontiguousarray(xy).ctypes.data, n,
                                        box.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        1.0, 32, nb.ctypes.data, cnt.ctypes.data, PREC_EXACT, 4,
                                        ctypes.byref(req))
    assert st == _lib.GDT_ERR_INVALID


def build_exact_audit(path, extra=()):
    subprocess.run(["g++", "-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-DGDT_AUDIT_EXACT",
                    *extra, f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(CSRC, "hostpath.cpp"),
                    os.path.join(ROOT, "tests", "cpp", "exact_audit.cpp"), "-o", path], check=True)
    return path


def run_exact_audit(exe, xy, lx, ly, cell, tmp_path):
    f = tmp_path / "pts.bin"
    with open(f, "wb") as fh:
        fh.write(np.int32(xy.shape[0]).tobytes())
        fh.write(np.array([lx, ly, cell], dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(xy, dtype=np.float64).tobytes())
    out = subprocess.run([exe, str(f)], capture_output=True, text=True, check=True).stdout.split()
    st, v = int(out[0]), [int(x) for x in out[1:]]
    return st, v[0::2], v[1::2]


def test_exact_mode_fp64_decisions_audited(tmp_path):
    """Every decision the exact routine takes from its FP64 value alone agrees
    with the exact sign of the predicate (the error bounds hold), on the
    near-degenerate lattices and generic points; with the bounds shrunk a
    million times the same audit finds disagreements, so it can fail."""
    exe = build_exact_audit(str(tmp_path / "audit"))
    tight = build_exact_audit(str(tmp_path / "audit_tight"), ["-DGDT_XTOL_SCALE=1e-6"])
    cases = [NEAR[k] for k in sorted(NEAR)]
    cases.append(((inputs.uniform_points(2000, 45.0, 45.0, seed=4), 45.0, 45.0), 1.0))
    wrong_tight = 0
    for (xy, lx, ly), cell in cases:
        st, taken, wrong = run_exact_audit(exe, xy, lx, ly, cell, tmp_path)
        assert st == 0 and min(taken) > 0
        assert wrong == [0, 0, 0], (taken, wrong)
        _, _, wt = run_exact_audit(tight, xy, lx, ly, cell, tmp_path)
        wrong_tight += sum(wt)
    assert wrong_tight > 0
