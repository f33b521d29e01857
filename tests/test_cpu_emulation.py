"""The device routines compiled for the host with g++ under AddressSanitizer
and UBSan and run on the CPU: the direct per-point routine (gdt_point.h), the
LDS-staged tile routine (gdt_tile.h, one emulated block per tile, deferred
points through the wavefront routine), the wavefront routine (gdt_wave.h,
64 emulated lanes and the shuffle reduction tree) and the star kernel's steps
(64 emulated lanes each searching from one candidate of the point's block,
the ring walked over their results).  FP64 and FP32-filtered
paths must reproduce the oracle's rings bit-exactly and be byte-identical to
each other, on uniform, near-degenerate (jittered lattice), rectangular, small,
fine-grid (cells much smaller than the spacing: circle-box extensions and
deferrals) and clustered (windows too full for LDS) inputs; the tile routine
for both compiled tile edges (32x32 and 16x16 cells)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from gpudelaunaytriangulation_amd import inputs, rings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_emul", "emul.cpp")
CSRC = os.path.join(ROOT, "gpudelaunaytriangulation_amd", "csrc")


# the LDS tile routine's two compiled tile edges: 32x32 cells (window budget
# 2048 points) and 16x16 (1024), as the library builds them
TILE_EDGES = {5: 2048, 4: 1024}


@pytest.fixture(scope="module", params=sorted(TILE_EDGES, reverse=True), ids=lambda t: f"ts{t}")
def emul(request, tmp_path_factory):
    ts = request.param
    exe = str(tmp_path_factory.mktemp("emul") / f"emul_ts{ts}")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-ffp-contract=off",
                    f"-DGDT_TILE_SHIFT={ts}", f"-DGDT_WIN_MAXPTS={TILE_EDGES[ts]}",
                    f"-I{CSRC}", SRC, "-o", exe], check=True)
    return exe


def run(exe, tmp, xy, lx, ly, tiled, prec, stride=32, mode=0, cell=1.0):
    n = xy.shape[0]
    pin, pout = os.path.join(tmp, "p.bin"), os.path.join(tmp, "o.bin")
    xy.astype(np.float64).tofile(pin)
    subprocess.run([exe, pin, str(n), repr(lx), repr(ly), repr(cell), str(tiled), str(prec),
                    str(stride), pout, str(mode)], check=True)
    raw = np.fromfile(pout, dtype=np.uint8)
    counts = raw[:4 * n].view(np.int32)
    nbrs = raw[4 * n:4 * n + 4 * n * stride].view(np.int32).reshape(n, stride)
    tail = raw[4 * n + 4 * n * stride:]
    npred, nfb = tail[:16].view(np.uint64)
    err = int(tail[16:20].view(np.int32)[0])
    return nbrs.copy(), counts.copy(), int(npred), int(nfb), err


def clustered():
    """2048 uniform points in a 64x64 box plus 2048 in a 3x3 square."""
    xy = inputs.uniform_points(4096, 64.0, 64.0, seed=11)
    xy[2048:] = 30.0 + xy[2048:] * (3.0 / 64.0)
    return xy


# name: (points, lx, ly, cell size)
CASES = {
    "uniform4096": (lambda: inputs.uniform_points(4096, 64.0, 64.0, seed=1), 64.0, 64.0, 1.0),
    "lattice64": (lambda: inputs.lattice_points(64, 64, 1e-5, seed=3), *inputs.lattice_box(64, 64),
                  1.0),
    "rect1000": (lambda: inputs.uniform_points(1000, 40.0, 25.0, seed=7), 40.0, 25.0, 1.0),
    "small300": (lambda: inputs.uniform_points(300, 17.0, 17.0, seed=9), 17.0, 17.0, 1.0),
    "tiny60": (lambda: inputs.uniform_points(60, 8.0, 8.0, seed=9), 8.0, 8.0, 1.0),
    "finecells": (lambda: inputs.uniform_points(2048, 45.0, 45.0, seed=5), 45.0, 45.0, 0.4),
    # 1.44 points per cell: the star steps' 7x7 blocks mostly exceed 64
    # entries (their 5x5 blocks), the 16x16-tile windows still fit
    "coarsecells": (lambda: inputs.uniform_points(3000, 55.0, 55.0, seed=6), 55.0, 55.0, 1.2),
    # a clump of 150 points in a 2x2 square: blocks beyond 64 entries even
    # at 5x5 (the star steps' sequential lane), in windows that still fit
    "bump": (lambda: np.concatenate([inputs.uniform_points(4096, 64.0, 64.0, seed=12),
                                     20.0 + 2.0 * inputs.uniform_points(150, 1.0, 1.0, seed=13)]),
             64.0, 64.0, 1.0),
    "clustered": (clustered, 64.0, 64.0, 1.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_emulated_kernels_match_oracle(emul, tmp_path, case):
    gen, lx, ly, cell = CASES[case]
    xy = gen()
    on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
    ref = rings.ring_hash(on, oc)
    outs = {}
    # (cell order, mode): direct routine in tile and row-major order, tile
    # routine in tile and Morton-in-tile order, wavefront routine in tile and
    # row-major order, the star kernel's steps in tile order
    for tiled, mode in ((1, 0), (0, 0), (1, 1), (2, 1), (1, 2), (0, 2), (1, 3)):
        for prec in (64, 32):
            nb, c, npred, nfb, err = run(emul, str(tmp_path), xy, lx, ly, tiled, prec, mode=mode,
                                         cell=cell)
            assert err == 0
            assert rings.ring_hash(nb, c) == ref, (prec, tiled, mode)
            if prec == 64 or mode == 3:
                assert nfb == 0
            elif case == "uniform4096" and mode == 1:
                # the LDS routine's filter decides all but a few per mille in
                # FP32 (about 1.4e-3 of the predicates on uniform input); v
                # and a are excluded by entry, not by their orientation
                assert nfb < 0.005 * npred, (tiled, nfb, npred)
            outs[(prec, tiled, mode)] = (nb, c)
        # identical decisions: FP32-filtered rows equal FP64 rows byte for byte
        n64, c64 = outs[(64, tiled, mode)]
        n32, c32 = outs[(32, tiled, mode)]
        assert np.array_equal(c64, c32)
        for i in range(len(c64)):
            assert np.array_equal(n64[i, :c64[i]], n32[i, :c64[i]])


@pytest.mark.parametrize("case", ["uniform4096", "lattice64", "finecells", "clustered"])
def test_emulated_filter_forced_to_fp64(tmp_path, case):
    """The FP32 filter taking no decision of its own (GDT_FILTER_FORCE_FP64: a
    certain FP32 verdict is never trusted): every predicate goes to the FP64
    fallback, and the rows equal the FP64 path's byte for byte, and the
    oracle's, in the direct and the tile routines."""
    exe = str(tmp_path / "emul_forced")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-ffp-contract=off",
                    "-DGDT_TILE_SHIFT=5", "-DGDT_WIN_MAXPTS=2048", "-DGDT_FILTER_FORCE_FP64",
                    f"-I{CSRC}", SRC, "-o", exe], check=True)
    gen, lx, ly, cell = CASES[case]
    xy = gen()
    on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
    ref = rings.ring_hash(on, oc)
    for tiled, mode in ((1, 0), (1, 1), (2, 1)):
        n64, c64, _, _, err = run(exe, str(tmp_path), xy, lx, ly, tiled, 64, mode=mode, cell=cell)
        assert err == 0
        n32, c32, npred, nfb, err = run(exe, str(tmp_path), xy, lx, ly, tiled, 32, mode=mode,
                                        cell=cell)
        assert err == 0
        # (the rest: entries of v and a scanned by the tile step, and the
        # nearest-neighbour search's first candidate, which no value decides;
        # on fine cells and clusters also the deferred points' FP64-only
        # wavefront routine, which takes all of the cluster's windows)
        least = {"uniform4096": 0.85, "lattice64": 0.85, "finecells": 0.2, "clustered": 0.0}
        assert nfb >= least[case] * npred, (tiled, mode, nfb, npred)
        assert rings.ring_hash(n32, c32) == ref, (tiled, mode)
        assert np.array_equal(c64, c32)
        for i in range(len(c64)):
            assert np.array_equal(n64[i, :c64[i]], n32[i, :c64[i]])


def _flip(nb, c, v, m):
    """Rings after flipping edge (v, b = ring_v[m]) of triangles (v, a, b) and
    (v, b, d) to edge (a, d)."""
    R = [list(nb[i, :c[i]]) for i in range(len(c))]
    cv = len(R[v])
    a, b, d = R[v][(m - 1) % cv], R[v][m], R[v][(m + 1) % cv]
    R[v].remove(b)
    R[b].remove(v)
    R[a].insert(R[a].index(v), d)       # around a: b, d, v
    R[d].insert(R[d].index(b), a)       # around d: v, a, b
    out = np.full_like(nb, -1)
    cnt = np.array([len(r) for r in R], dtype=c.dtype)
    for i, r in enumerate(R):
        out[i, :len(r)] = r
    return out, cnt


# exact square lattices: every square's four points are cocircular
TIES = {"square16": (16, 16, 1.0), "rect12x20": (12, 20, 1.0), "half16": (16, 16, 0.5)}


@pytest.mark.parametrize("case", sorted(TIES))
def test_emulated_kernels_exact_ties(emul, tmp_path, case):
    """Exact ties (cocircular squares, equidistant neighbours) are decided by
    the symbolic perturbation |p|^2 + eps^(1 + caller index): every path and
    precision gives the same rings, and they are the Delaunay triangulation of
    the perturbed points (checked with exact integer determinants and the
    perturbation's leading term, not with the kernels' tie code)."""
    from exact_delaunay import SCALE, perturbed_violations, ring_inconsistencies, square_lattice
    nx, ny, sp = TIES[case]
    P = square_lattice(nx, ny, sp)
    xy, lx, ly = P / SCALE, nx * sp, ny * sp
    ref = None
    for tiled, mode in ((1, 0), (0, 0), (1, 1), (2, 1), (1, 2), (0, 2), (1, 3)):
        for prec in (64, 32):
            nb, c, _, _, err = run(emul, str(tmp_path), xy, lx, ly, tiled, prec, mode=mode)
            assert err == 0
            if ref is None:
                ref = rings.ring_hash(nb, c)
                assert int(c.sum()) == 6 * len(c)
                assert ring_inconsistencies(nb, c) == 0
                assert perturbed_violations(P, lx, ly, nb, c) == 0
                # the other diagonal of a square: still a triangulation with
                # empty (unperturbed) circles, but not the perturbed one's
                # a diagonal: neither the same row nor the same column
                v, m = next((v, m) for v in range(len(c)) for m in range(c[v])
                            if (P[nb[v, m]] != P[v]).all())
                fn, fc = _flip(nb, c, v, m)
                assert ring_inconsistencies(fn, fc) == 0
                assert perturbed_violations(P, lx, ly, fn, fc) > 0
            assert rings.ring_hash(nb, c) == ref, (tiled, mode, prec)


def test_tie_rule_matches_perturbed_determinant(tmp_path):
    """gdt_incircle_tie against the leading term of the perturbed in-circle
    determinant (exact integers, tests/exact_delaunay.py), for every ordered
    choice of v, a, b, c among the 12 integer points of the circle
    x^2 + y^2 = 25 with b and c left of v->a, and all 24 ways to rank their
    caller indices: all four branches of the rule (which point is raised
    most) meet every configuration of the circle."""
    from itertools import permutations

    from exact_delaunay import _sos_sign
    exe = str(tmp_path / "tie_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", f"-I{CSRC}",
                    os.path.join(ROOT, "tests", "cpu_emul", "tie_check.cpp"), "-o", exe], check=True)
    circ = [(x, y) for x in range(-5, 6) for y in range(-5, 6) if x * x + y * y == 25]
    assert len(circ) == 12
    lines = []
    for v in circ:
        others = [(x - v[0], y - v[1]) for (x, y) in circ if (x, y) != v]
        for A, B, C in permutations(others, 3):
            if A[0] * B[1] - A[1] * B[0] <= 0 or A[0] * C[1] - A[1] * C[0] <= 0:
                continue
            p, q, r = (-C[0], -C[1]), (A[0] - C[0], A[1] - C[1]), (B[0] - C[0], B[1] - C[1])
            cp = q[0] * r[1] - q[1] * r[0]
            cq = -(p[0] * r[1] - p[1] * r[0])
            cr = p[0] * q[1] - p[1] * q[0]
            for pv, pa, pb, pc in permutations(range(4)):
                lines.append((pv, pa, pb, pc, A, B, C, cp, cq, cr))
    # determinant rows v - c, a - c, b - c (cofactors cp, cq, cr of the lifted
    # column); the tested point c carries -(cp + cq + cr); > 0: c inside
    exp = []
    for (pv, pa, pb, pc, A, B, C, cp, cq, cr) in lines:
        s = _sos_sign(0, [cp, cq, cr, -(cp + cq + cr)], [pv, pa, pb, pc])
        assert s != 0
        exp.append(1 if s > 0 else 0)
    inp = "".join(f"ic {pv} {pa} {pb} {pc} {A[0]} {A[1]} {B[0]} {B[1]} {C[0]} {C[1]}\n"
                  for (pv, pa, pb, pc, A, B, C, *_r) in lines)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    assert len(got) == len(exp) > 1000
    bad = [i for i in range(len(exp)) if got[i] != exp[i]]
    assert not bad, (len(bad), lines[bad[0]][:7])
