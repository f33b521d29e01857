"""General (triclinic) periodic boxes on the CPU: the oracle's general-box
restatement and the library's CPU mode (the kernels' own routines compiled
for the host), no GPU.

The reference has general boxes in periodicBoundaries (setGeneral, general
minDist and putInBoxReal, periodicBoundaries.h:86-178) and a square-only cell
list (cellListGPU.cpp:97-117).  Pinned here:

* the oracle (cells in unit-cell coordinates, minDist's general branch)
  equals exact integer brute force on sheared integer lattices and Qhull on
  nine-sheet padded copies of sheared boxes; a transposed box matrix (a
  plausible slip) is caught;
* the CPU mode (the device routines: cells of the Lagrange-Gauss reduced
  basis, parallelogram cell displacements, coverage by cell heights) equals
  Qhull and the oracle, for FP64 and FP32-filtered requests, for boxes whose
  basis is not reduced (tilt above one), left-handed and negative-diagonal
  matrices; test-and-repair after a move equals a fresh triangulation;
* a rectangular matrix passed to the general calls gives the rectangular
  calls' rings bit for bit;
* Voronoi vertices in a general box are the circumcentres of minimum-image
  neighbours; the Voronoi energy of the oracle in a sheared box agrees with
  the energy of the same tessellation in an equivalent rectangle;
* PeriodicBoundaries(a, b, c, d): Trans/invTrans are inverse, minDist is the
  shortest image of short displacements, putInBoxReal lands in the unit cell.
"""
import ctypes

import numpy as np
import pytest

import oracle
from exact_delaunay import SCALE, lattice_points, rings as exact_rings
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, _lib, inputs, rings
from qhull_periodic import periodic_rings_box

torch = pytest.importorskip("torch")

from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries  # noqa: E402


def canon_lists(ring_lists):
    counts = np.array([len(r) for r in ring_lists], dtype=np.int32)
    nb = np.full((len(ring_lists), counts.max()), -1, dtype=np.int32)
    for i, r in enumerate(ring_lists):
        nb[i, :len(r)] = r
    return counts, rings.canonical(nb, counts)


def canon(nb, c):
    return rings.canonical(nb, c)


# ---- oracle ----------------------------------------------------------------

EXACT_BOXES = [((7, 2, 0, 7), 48, 1), ((8, -3, 1, 7), 60, 2), ((9, 4, 2, 8), 70, 3)]


@pytest.mark.parametrize("box,n,seed", EXACT_BOXES)
def test_oracle_general_box_equals_exact_bruteforce(box, n, seed):
    P = lattice_points(n, box, seed)
    xy = P.astype(np.float64) / SCALE
    ex_counts, ex_can = canon_lists(exact_rings(P, 0, 0, D=4.0, box_cells=box, images=2))
    nb, c = oracle.triangulate_box(xy, np.array(box, float), 1.0, 16)
    assert np.array_equal(c, ex_counts)
    assert np.array_equal(canon(nb, c), ex_can)


def test_oracle_transposed_box_is_caught():
    """x12 and x21 swapped describes another lattice: the rings differ."""
    box = (8, -3, 1, 7)
    P = lattice_points(60, box, 2)
    xy = P.astype(np.float64) / SCALE
    _, ex_can = canon_lists(exact_rings(P, 0, 0, D=4.0, box_cells=box, images=2))
    nb, c = oracle.triangulate_box(xy, np.array([8.0, 1.0, -3.0, 7.0]), 1.0, 16)
    assert c.shape != ex_can.shape or not np.array_equal(canon(nb, c), ex_can)


QHULL_BOXES = [(45.0, 12.0, 0.0, 44.4), (40.0, -15.3, 3.1, 50.0), (52.0, 0.0, 9.5, 38.0)]


@pytest.mark.parametrize("box", QHULL_BOXES)
def test_oracle_general_box_equals_qhull(box):
    xy = inputs.general_points(2000, box, seed=21)
    nb, c = oracle.triangulate_box(xy, box, 1.0, 16)
    qn, qc = periodic_rings_box(xy, box, margin=3.0)
    assert np.array_equal(c, qc)
    assert np.array_equal(canon(nb, c), canon(qn, qc))
    rings.check_torus_invariants(nb, c)


def test_oracle_diagonal_box_is_rectangular():
    xy = inputs.uniform_points(1500, 40.0, 30.0, seed=4)
    nb, c = oracle.triangulate(xy, 40.0, 30.0, 1.0, 16)
    gb, gc = oracle.triangulate_box(xy, (40.0, 0.0, 0.0, 30.0), 1.0, 16)
    assert np.array_equal(nb, gb) and np.array_equal(c, gc)


def test_oracle_general_test_and_repair_equals_fresh():
    box = QHULL_BOXES[1]
    xy = inputs.general_points(3000, box, seed=22)
    nb, c = oracle.triangulate_box(xy, box, 1.0, 16)
    rng = np.random.default_rng(5)
    moved = inputs.wrap_into_general_box(xy + rng.uniform(-0.08, 0.08, xy.shape), box)
    rn, rc, nflag = oracle.test_and_repair_box(moved, box, nb, c)
    fn, fc = oracle.triangulate_box(moved, box, 1.0, 16)
    assert nflag > 0
    assert rings.ring_hash(rn, rc) == rings.ring_hash(fn, fc)


def test_oracle_voronoi_energy_is_box_shape_independent():
    """The same periodic point set described by a sheared basis of a square
    lattice ((L, L), (0, L) spans the square L x L torus): areas, perimeters
    and gradients equal those of the rectangular description."""
    L = 40.0
    xy = inputs.uniform_points(1600, L, L, seed=23)
    nb, c = oracle.triangulate(xy, L, L, 1.0, 16)
    E, A, P, G = oracle.voronoi_energy(xy, L, L, nb, c)
    Eg, Ag, Pg, Gg = oracle.voronoi_energy_box(xy, (L, 0.0, L, L), nb, c)
    assert np.allclose(A, Ag, rtol=1e-10, atol=1e-10)
    assert np.allclose(P, Pg, rtol=1e-10, atol=1e-10)
    assert np.allclose(G, Gg, rtol=1e-8, atol=1e-8)
    assert abs(E - Eg) <= 1e-8 * abs(E)


# ---- CPU mode (the device routines on the host) ---------------------------

def host_tri_box(xy, box, stride=32, prec=PREC_FP64, threads=4, cell=1.0):
    L = _lib.load()
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    nb = np.full((n, stride), -1, dtype=np.int32)
    cnt = np.zeros(n, dtype=np.int32)
    req = ctypes.c_int(0)
    b = (ctypes.c_double * 4)(*[float(v) for v in box])
    st = L.gdt_host_triangulate_general(xy.ctypes.data, n, b, cell, stride, nb.ctypes.data,
                                        cnt.ctypes.data, prec, threads, ctypes.byref(req))
    return st, nb, cnt, req.value


HOST_BOXES = {
    "shear": (45.0, 12.0, 0.0, 44.4),
    "general": (40.0, -15.3, 3.1, 50.0),
    "unreduced": (30.0, 75.0, 0.0, 66.0),      # tilt 75/66 > 1: reduced basis differs
    "lefthanded": (0.0, 48.0, 42.0, 6.0),      # det < 0
    "negdiag": (-40.0, 0.0, 0.0, 50.0),
}


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
@pytest.mark.parametrize("name", list(HOST_BOXES))
def test_host_general_box_equals_qhull_and_oracle(name, prec):
    box = HOST_BOXES[name]
    xy = inputs.general_points(2000, box, seed=31)
    st, nb, c, _ = host_tri_box(xy, box, prec=prec)
    assert st == _lib.GDT_OK
    qn, qc = periodic_rings_box(xy, box, margin=3.0)
    assert np.array_equal(canon(nb, c), canon(qn, qc))
    if name != "unreduced":     # the oracle follows minDist's unit-cell rounding of the given basis
        on, oc = oracle.triangulate_box(xy, box, 1.0, 16)
        assert np.array_equal(canon(nb, c), canon(on, oc))


def test_host_general_calls_on_a_rectangle_are_the_rectangular_calls():
    L = _lib.load()
    xy = inputs.uniform_points(3000, 60.0, 50.0, seed=8)
    st, nb, c, _ = host_tri_box(xy, (60.0, 0.0, 0.0, 50.0))
    nb2 = np.full_like(nb, -1)
    c2 = np.zeros_like(c)
    req = ctypes.c_int(0)
    st2 = L.gdt_host_triangulate(xy.ctypes.data, 3000, 60.0, 50.0, 1.0, 32, nb2.ctypes.data,
                                 c2.ctypes.data, PREC_FP64, 4, ctypes.byref(req))
    assert st == st2 == _lib.GDT_OK
    assert np.array_equal(nb, nb2) and np.array_equal(c, c2)


def test_host_general_box_threads_agree():
    box = HOST_BOXES["general"]
    xy = inputs.general_points(2500, box, seed=32)
    _, nb1, c1, _ = host_tri_box(xy, box, threads=1)
    _, nb8, c8, _ = host_tri_box(xy, box, threads=8)
    assert np.array_equal(nb1, nb8) and np.array_equal(c1, c8)


def test_python_cpu_mode_general_box_test_and_repair():
    """DelaunayGPU with PeriodicBoundaries(a, b, c, d) in CPU mode: global
    triangulation from MaxSize 4, a move, test-and-repair equals a fresh
    oracle triangulation of the moved points."""
    box = HOST_BOXES["shear"]
    n = 3000
    xy = inputs.general_points(n, box, seed=33)
    dg = DelaunayGPU(n, 4, 1.0, PeriodicBoundaries(*box))
    dg.setGPUcompute(False)
    P = torch.tensor(xy)
    T = torch.zeros(n * 4, dtype=torch.int32)
    C = torch.zeros(n, dtype=torch.int32)
    dg.globalDelaunayTriangulation(P, T, C)
    on, oc = oracle.triangulate_box(xy, box, 1.0, 16)
    assert dg.MaxSize > 4
    assert np.array_equal(canon(T.numpy().reshape(n, dg.MaxSize), C.numpy()), canon(on, oc))
    moved = inputs.wrap_into_general_box(xy + 0.04 * inputs.gaussian(34, 2 * n).reshape(n, 2), box)
    dg.testAndRepairDelaunayTriangulation(torch.tensor(moved), T, C)
    fn, fc = oracle.triangulate_box(moved, box, 1.0, 16)
    assert dg.last_repaired > 0
    assert np.array_equal(canon(T.numpy().reshape(n, dg.MaxSize), C.numpy()), canon(fn, fc))


def test_host_general_voronoi_vertices():
    L = _lib.load()
    box = HOST_BOXES["general"]
    X = np.array(box).reshape(2, 2)
    Xi = np.linalg.inv(X)
    n = 1500
    xy = inputs.general_points(n, box, seed=35)
    st, nb, c, _ = host_tri_box(xy, box)
    assert st == _lib.GDT_OK
    out = np.zeros((n, 32, 2))
    b = (ctypes.c_double * 4)(*box)
    assert L.gdt_host_voronoi_vertices_general(xy.ctypes.data, n, b, nb.ctypes.data, c.ctypes.data,
                                               32, out.ctypes.data, 4) == _lib.GDT_OK
    for i in range(0, n, 7):
        k = c[i]
        d = xy[nb[i, :k]] - xy[i]
        w = d @ Xi.T
        d = (w - np.round(w)) @ X.T
        a, bb = d, np.roll(d, -1, axis=0)
        an, bn = (a * a).sum(1), (bb * bb).sum(1)
        den = 2.0 * (a[:, 0] * bb[:, 1] - a[:, 1] * bb[:, 0])
        ref = np.stack([(an * bb[:, 1] - bn * a[:, 1]) / den, (bn * a[:, 0] - an * bb[:, 0]) / den], 1)
        assert np.allclose(out[i, :k], ref, rtol=1e-10, atol=1e-10)


def test_host_small_general_box_finishes_with_the_wavefront_routine(tmp_path):
    """6 x 6 cells of a sheared box: the per-point routine (host build of the
    device routine, tests/cpu_emul/emul.cpp) cannot certify some searches
    within the grid's search radius; the CPU mode hands those points to the
    wavefront routine, as the device's row-major kernels do, and the rings
    equal Qhull."""
    import subprocess
    from test_filter_audit import SRC, CSRC
    exe = str(tmp_path / "emul")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{CSRC}", SRC, "-o", exe],
                   check=True)
    box = (7.0, 2.0, 0.5, 7.0)
    n = 50
    xy = inputs.general_points(n, box, seed=40)
    pin, pout = str(tmp_path / "p.bin"), str(tmp_path / "o.bin")
    xy.tofile(pin)
    errs = []
    for mode in (0, 2):     # direct per-point routine, wavefront routine
        subprocess.run([exe, pin, str(n), "box:" + ",".join(repr(v) for v in box), "0", "1.0", "0",
                        "64", "32", pout, str(mode)], check=True)
        errs.append(int(np.fromfile(pout, dtype=np.int32)[-1]))
    assert errs == [1, 0]       # ERR_SEARCH from the direct routine only
    st, nb, c, _ = host_tri_box(xy, box)
    assert st == _lib.GDT_OK
    qn, qc = periodic_rings_box(xy, box, margin=2.0)
    assert np.array_equal(canon(nb, c), canon(qn, qc))


def test_host_general_box_rejects_singular_matrix():
    xy = inputs.uniform_points(100, 10.0, 10.0, seed=1)
    st, *_ = host_tri_box(xy, (10.0, 20.0, 5.0, 10.0))     # det = 0
    assert st == _lib.GDT_ERR_INVALID
    st, *_ = host_tri_box(xy, (10.0, np.nan, 0.0, 10.0))
    assert st == _lib.GDT_ERR_INVALID


# ---- the box class ---------------------------------------------------------

def test_periodic_boundaries_general_box():
    B = PeriodicBoundaries(10.0, 3.0, 1.0, 8.0)
    assert not B.isBoxSquare() and not B.is_rectangle()
    rng = np.random.default_rng(2)
    p = rng.uniform(-30, 30, size=(200, 2))
    assert np.allclose(B.Trans(B.invTrans(p)), p, rtol=1e-13, atol=1e-12)
    q = B.putInBoxReal(p)
    u = B.invTrans(q)
    assert np.all(u >= -1e-12) and np.all(u < 1 + 1e-12)
    # q differs from p by whole lattice vectors
    k = B.invTrans(p - q)
    assert np.allclose(k, np.round(k), atol=1e-9)
    # minDist of short displacements is the shortest image (brute force over images)
    X = np.array([[10.0, 3.0], [1.0, 8.0]])
    for _ in range(100):
        a = rng.uniform(0, 10, 2)
        d = rng.uniform(-2, 2, 2)
        b = B.putInBoxReal(a + d)
        got = B.minDist(b, a)
        imgs = [b - a + i * X[:, 0] + j * X[:, 1] for i in range(-3, 4) for j in range(-3, 4)]
        best = min(imgs, key=lambda v: float(v @ v))
        assert np.allclose(got, best, atol=1e-12)
    with pytest.raises(ValueError):
        PeriodicBoundaries(1.0, 2.0, 2.0, 4.0)
    R = PeriodicBoundaries(5.0, 0.0, 0.0, 7.0)
    assert R.is_rectangle() and not R.isBoxSquare()     # setGeneral keeps isSquare false
