"""CPU mode of the library (the reference's DelaunayGPU::setGPUcompute(false)
with setOMPthreads, DelaunayGPU.h:48-50, DelaunayGPU.cpp:159-343, 567-706):
the kernels' own per-point routines compiled for the host (gdt_host_* in
libgpudelaunay.so), checked against the oracle on this machine, no GPU.

* global triangulation, FP64 and FP32-filtered, on uniform, jittered-lattice,
  rectangular, small, fine-grid and clustered inputs: rings bit-identical to
  the oracle's (canonical), for 1 and 8 OpenMP threads;
* the 2^20-point uniform workload equals the committed golden hash;
* test-and-repair after moving 5 % of the points equals a fresh oracle
  triangulation; a valid triangulation flags nothing; rings that are not a
  triangulation give GDT_ERR_EULER;
* the Python DelaunayGPU in CPU mode grows MaxSize from 4 and fills its
  profiler; the host cell list is the closed-form binning in index order;
  host Voronoi vertices are the circumcentres of consecutive ring neighbours.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, _lib, inputs, rings

torch = pytest.importorskip("torch")

from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def host_tri(xy, lx, ly, stride=32, prec=PREC_FP64, threads=4, cell=1.0):
    L = _lib.load()
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    nb = np.full((n, stride), -1, dtype=np.int32)
    cnt = np.zeros(n, dtype=np.int32)
    req = ctypes.c_int(0)
    st = L.gdt_host_triangulate(xy.ctypes.data, n, lx, ly, cell, stride, nb.ctypes.data,
                                cnt.ctypes.data, prec, threads, ctypes.byref(req))
    return st, nb, cnt, req.value


def clustered():
    xy = inputs.uniform_points(4096, 64.0, 64.0, seed=11)
    xy[2048:] = 30.0 + xy[2048:] * (3.0 / 64.0)
    return xy


CASES = {
    "uniform4096": (lambda: inputs.uniform_points(4096, 64.0, 64.0, seed=1), 64.0, 64.0, 1.0),
    "lattice64": (lambda: inputs.lattice_points(64, 64, 1e-5, seed=3), *inputs.lattice_box(64, 64),
                  1.0),
    "rect1000": (lambda: inputs.uniform_points(1000, 40.0, 25.0, seed=7), 40.0, 25.0, 1.0),
    "small300": (lambda: inputs.uniform_points(300, 17.0, 17.0, seed=9), 17.0, 17.0, 1.0),
    "finecells": (lambda: inputs.uniform_points(2048, 45.0, 45.0, seed=5), 45.0, 45.0, 0.4),
    "clustered": (clustered, 64.0, 64.0, 1.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_host_triangulation_matches_oracle(case, prec):
    gen, lx, ly, cell = CASES[case]
    xy = gen()
    on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
    want = rings.ring_hash(on, oc)
    for threads in (1, 8):
        st, nb, cnt, req = host_tri(xy, lx, ly, prec=prec, threads=threads, cell=cell)
        assert st == _lib.GDT_OK, _lib.status_string(st)
        assert req == int(oc.max())
        assert rings.ring_hash(nb, cnt) == want, f"{case}, {threads} threads"


def test_host_triangulation_c2_golden_hash():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        g = json.load(fh)["workloads"]["c2"]
    xy = inputs.uniform_points(1 << 20, 1024.0, 1024.0, seed=2)
    assert inputs.sha256(xy) == g["points_sha256"]
    st, nb, cnt, _ = host_tri(xy, 1024.0, 1024.0, stride=24, threads=8)
    assert st == _lib.GDT_OK
    assert int(cnt.sum()) == 6 * xy.shape[0]
    assert rings.ring_hash(nb, cnt) == g["ring_sha256"]


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_host_grid_smaller_than_the_halo(prec):
    """6x6 cells: the per-point search reaches the grid's search radius and
    the wavefront routine (the device's deferred path) finishes those points;
    the rings equal scipy's Qhull."""
    from qhull_periodic import periodic_rings
    xy = inputs.uniform_points(40, 6.0, 6.0, seed=43)
    qn, qc = periodic_rings(xy, 6.0, 6.0)
    st, nb, cnt, _ = host_tri(xy, 6.0, 6.0, prec=prec)
    assert st == _lib.GDT_OK, _lib.status_string(st)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(qn, qc)


def test_host_overflow_reports_required():
    xy = inputs.uniform_points(4096, 64.0, 64.0, seed=1)
    st, _, cnt, req = host_tri(xy, 64.0, 64.0, stride=5)
    assert st == _lib.GDT_OVERFLOW and req == int(cnt.max()) > 5


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_host_test_and_repair_equals_fresh_triangulation(prec):
    L = _lib.load()
    n, Lb = 4096, 64.0
    xy = inputs.uniform_points(n, Lb, Lb, seed=1)
    st, nb, cnt, _ = host_tri(xy, Lb, Lb, stride=20, prec=prec)
    assert st == _lib.GDT_OK
    nf, req = ctypes.c_int(-1), ctypes.c_int(0)
    # a valid triangulation: nothing flagged, rows unchanged
    nb0, cnt0 = nb.copy(), cnt.copy()
    st = L.gdt_host_test_and_repair(xy.ctypes.data, n, Lb, Lb, 1.0, 20, nb.ctypes.data,
                                    cnt.ctypes.data, prec, 4, ctypes.byref(nf), ctypes.byref(req))
    assert st == _lib.GDT_OK and nf.value == 0
    assert np.array_equal(nb, nb0) and np.array_equal(cnt, cnt0)
    rng = np.random.default_rng(0)
    sel = rng.choice(n, size=n // 20, replace=False)
    moved = xy.copy()
    moved[sel] += rng.uniform(-0.1, 0.1, size=(sel.size, 2))
    moved = inputs.wrap_into_box(moved, Lb, Lb)
    st = L.gdt_host_test_and_repair(moved.ctypes.data, n, Lb, Lb, 1.0, 20, nb.ctypes.data,
                                    cnt.ctypes.data, prec, 4, ctypes.byref(nf), ctypes.byref(req))
    assert st == _lib.GDT_OK, _lib.status_string(st)
    assert nf.value > 0
    fn, fc = oracle.triangulate(moved, Lb, Lb, 1.0, 16)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(fn, fc)


def test_host_test_and_repair_rejects_non_triangulation():
    L = _lib.load()
    n, Lb = 1024, 32.0
    xy = inputs.uniform_points(n, Lb, Lb, seed=3)
    st, nb, cnt, _ = host_tri(xy, Lb, Lb, stride=20)
    cnt[17] -= 1                       # drop one neighbour: 6N and 2N both fail
    nf, req = ctypes.c_int(0), ctypes.c_int(0)
    st = L.gdt_host_test_and_repair(xy.ctypes.data, n, Lb, Lb, 1.0, 20, nb.ctypes.data,
                                    cnt.ctypes.data, PREC_FP64, 2, ctypes.byref(nf),
                                    ctypes.byref(req))
    assert st == _lib.GDT_ERR_EULER


def test_python_cpu_mode_and_profiler():
    n, Lb = 4096, 64.0
    xy = inputs.uniform_points(n, Lb, Lb, seed=1)
    dg = DelaunayGPU(n, 4, 1.0, PeriodicBoundaries(Lb, Lb))
    dg.setGPUcompute(False)
    dg.setOMPthreads(4)
    pts = torch.from_numpy(xy.copy())
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32)
    num = torch.zeros(n, dtype=torch.int32)
    dg.globalDelaunayTriangulation(pts, tri, num)
    assert dg.MaxSize > 4 and tri.numel() == n * dg.MaxSize
    on, oc = oracle.triangulate(xy, Lb, Lb, 1.0, 16)
    assert rings.ring_hash(tri.numpy().reshape(n, dg.MaxSize), num.numpy()) == \
        rings.ring_hash(on, oc)
    moved = inputs.wrap_into_box(xy + 0.02 * inputs.gaussian(8, 2 * n).reshape(n, 2), Lb, Lb)
    dg.testAndRepairDelaunayTriangulation(torch.from_numpy(moved), tri, num)
    fn, fc = oracle.triangulate(moved, Lb, Lb, 1.0, 16)
    assert rings.ring_hash(tri.numpy().reshape(n, dg.MaxSize), num.numpy()) == \
        rings.ring_hash(fn, fc)
    assert dg.last_repaired > 0
    for name in ("initialization", "globalTriangulation", "testAndRepair"):
        assert dg.prof.functionCalls[dg.prof.processMap[name]] >= 1
        assert dg.prof.timing(name) > 0.0
    assert dg.prof.timing("no such region") == 0.0
    # CPU mode takes CPU tensors only (a tensor on any other device is refused)
    with pytest.raises(ValueError):
        dg.globalDelaunayTriangulation(pts.to("meta"), tri, num)


def test_host_cell_list_is_closed_form_binning():
    n, lx, ly = 3000, 40.0, 30.0
    xy = inputs.uniform_points(n, lx, ly, seed=12)
    L = _lib.load()
    xs, ys, mo = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.gdt_host_celllist(xy.ctypes.data, n, lx, ly, 1.0, None, None, 0, ctypes.byref(xs),
                               ctypes.byref(ys), ctypes.byref(mo)) == _lib.GDT_OK
    assert (xs.value, ys.value) == (40, 30)
    nmax = mo.value
    sizes = np.zeros(xs.value * ys.value, dtype=np.uint32)
    idxs = np.zeros(xs.value * ys.value * nmax, dtype=np.int32)
    assert L.gdt_host_celllist(xy.ctypes.data, n, lx, ly, 1.0, sizes.ctypes.data, idxs.ctypes.data,
                               nmax - 1, ctypes.byref(xs), ctypes.byref(ys),
                               ctypes.byref(mo)) == _lib.GDT_OVERFLOW
    assert L.gdt_host_celllist(xy.ctypes.data, n, lx, ly, 1.0, sizes.ctypes.data, idxs.ctypes.data,
                               nmax, ctypes.byref(xs), ctypes.byref(ys),
                               ctypes.byref(mo)) == _lib.GDT_OK
    cx = np.floor(xy[:, 0] / (lx / 40)).astype(int)
    cy = np.floor(xy[:, 1] / (ly / 30)).astype(int)
    bucket = cx + 40 * cy
    assert np.array_equal(sizes, np.bincount(bucket, minlength=1200))
    idxs = idxs.reshape(1200, nmax)
    for b in range(1200):
        assert np.array_equal(idxs[b, :sizes[b]], np.flatnonzero(bucket == b))


def test_host_voronoi_vertices_are_circumcentres():
    n, Lb = 1024, 32.0
    xy = inputs.uniform_points(n, Lb, Lb, seed=4)
    st, nb, cnt, _ = host_tri(xy, Lb, Lb, stride=16)
    out = np.zeros((n, 16, 2))
    L = _lib.load()
    assert L.gdt_host_voronoi_vertices(xy.ctypes.data, n, Lb, Lb, nb.ctypes.data, cnt.ctypes.data,
                                       16, out.ctypes.data, 4) == _lib.GDT_OK
    for i in range(0, n, 37):
        k = cnt[i]
        d = xy[nb[i, :k]] - xy[i]
        d -= Lb * np.round(d / Lb)
        a, b = d, np.roll(d, -1, axis=0)
        D = 2.0 * (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
        aa, bb = (a * a).sum(1), (b * b).sum(1)
        c = np.stack([(aa * b[:, 1] - bb * a[:, 1]) / D, (bb * a[:, 0] - aa * b[:, 0]) / D], 1)
        np.testing.assert_allclose(out[i, :k], c, rtol=1e-12, atol=1e-12)
        # each vertex is equidistant from the point and its two neighbours
        r0 = np.hypot(*c.T)
        np.testing.assert_allclose(np.hypot(*(c - a).T), r0, rtol=1e-10)
