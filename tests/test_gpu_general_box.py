"""General (triclinic) periodic boxes on the device, through the C ABI
(gdt_*_general via DelaunayGPU / BatchedDelaunayGPU / voronoi with
PeriodicBoundaries(a, b, c, d) or box matrices):

* global triangulation in sheared, unreduced (tilt above one), left-handed
  and negative-diagonal boxes: rings equal the oracle's general-box
  restatement and Qhull on a nine-sheet padded copy, and the device rows are
  bit-identical to the CPU mode's (the same routines on the host); an FP32
  request gives the FP64 rows (general boxes evaluate in FP64);
* a 262,144-point sheared box (the size-independent checks: the CPU mode's
  rows, Euler counts, torus invariants);
* test-and-repair after a move equals a fresh oracle triangulation;
* a batch mixing rectangles and general boxes equals each set alone;
* the Brownian loop in general boxes: positions bit-identical to the numpy
  restatement (counter noise, wrap by whole lattice vectors), rings equal to
  the oracle's triangulation of the final positions, graph = eager;
* Voronoi vertices (minimum images by rounded unit-cell coordinates) and the
  Voronoi energy and gradient against the oracle's general-box energy;
* a box of a few cells: points the row-major kernels cannot certify go to
  the wavefront kernel, as in the CPU mode.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, _lib, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import (BatchedDelaunayGPU, Context,  # noqa: E402
                                                   DelaunayGPU, PeriodicBoundaries)
from gpudelaunaytriangulation_amd.voronoi import voronoi_forward, voronoi_adjoint  # noqa: E402
from qhull_periodic import periodic_rings_box  # noqa: E402

BOXES = {
    "shear": (45.0, 12.0, 0.0, 44.4),
    "general": (40.0, -15.3, 3.1, 50.0),
    "unreduced": (30.0, 75.0, 0.0, 66.0),
    "lefthanded": (0.0, 48.0, 42.0, 6.0),
    "negdiag": (-40.0, 0.0, 0.0, 50.0),
}


def canon(nb, c):
    return rings.canonical(nb, c)


def same_rows(nb, c, hn, hc):
    """Identical counts and identical ring entries (slots past the count are
    not part of a row)."""
    if not np.array_equal(c, hc):
        return False
    live = np.arange(nb.shape[1])[None, :] < c[:, None]
    return bool(np.array_equal(np.where(live, nb, -1), np.where(live, hn, -1)))


def gpu_tri_box(xy, box, prec=PREC_FP64, maxn=16, cell=1.0):
    n = xy.shape[0]
    dg = DelaunayGPU(n, maxn, cell, PeriodicBoundaries(*box))
    dg.setPrecision(prec)
    P = torch.from_numpy(np.ascontiguousarray(xy)).cuda()
    T = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    C = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(P, T, C)
    torch.cuda.synchronize()
    return T.cpu().numpy().reshape(n, dg.MaxSize), C.cpu().numpy(), dg, (P, T, C)


def host_tri_box(xy, box, stride, cell=1.0):
    L = _lib.load()
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    nb = np.full((n, stride), -1, dtype=np.int32)
    cnt = np.zeros(n, dtype=np.int32)
    req = ctypes.c_int(0)
    b = (ctypes.c_double * 4)(*[float(v) for v in box])
    st = L.gdt_host_triangulate_general(xy.ctypes.data, n, b, cell, stride, nb.ctypes.data,
                                        cnt.ctypes.data, PREC_FP64, 8, ctypes.byref(req))
    assert st == _lib.GDT_OK
    return nb, cnt


@pytest.mark.parametrize("name", list(BOXES))
def test_general_box_equals_oracle_qhull_and_cpu_mode(name):
    box = BOXES[name]
    xy = inputs.general_points(4000, box, seed=41)
    print(f"[general box] {name}", flush=True)
    nb, c, dg, _ = gpu_tri_box(xy, box)
    assert int(c.sum()) == 6 * 4000
    qn, qc = periodic_rings_box(xy, box, margin=3.0)
    assert np.array_equal(canon(nb, c), canon(qn, qc))
    if name != "unreduced":     # the oracle rounds unit-cell coordinates of the given basis
        on, oc = oracle.triangulate_box(xy, box, 1.0, 16)
        assert np.array_equal(canon(nb, c), canon(on, oc))
    hn, hc = host_tri_box(xy, box, dg.MaxSize)
    assert same_rows(nb, c, hn, hc)


def test_fp32_request_gives_fp64_rows():
    box = BOXES["general"]
    xy = inputs.general_points(6000, box, seed=42)
    n64, c64, _, _ = gpu_tri_box(xy, box, PREC_FP64)
    n32, c32, dg, _ = gpu_tri_box(xy, box, PREC_FP32F)
    assert np.array_equal(n64, n32) and np.array_equal(c64, c32)
    assert dg.stats()["fallbacks"] == 0


def test_large_sheared_box_size_independent_checks():
    box = (512.0, 160.0, 0.0, 512.0)
    n = 1 << 18
    xy = inputs.general_points(n, box, seed=43)
    print("[general box] 262144 points", flush=True)
    nb, c, dg, _ = gpu_tri_box(xy, box)
    assert int(c.sum()) == 6 * n
    rings.check_torus_invariants(nb, c)
    hn, hc = host_tri_box(xy, box, dg.MaxSize)
    assert same_rows(nb, c, hn, hc)


def test_general_box_test_and_repair_equals_fresh():
    box = BOXES["shear"]
    n = 5000
    xy = inputs.general_points(n, box, seed=44)
    nb, c, dg, (P, T, C) = gpu_tri_box(xy, box)
    moved = inputs.wrap_into_general_box(xy + 0.05 * inputs.gaussian(45, 2 * n).reshape(n, 2), box)
    P.copy_(torch.from_numpy(moved))
    dg.testAndRepairDelaunayTriangulation(P, T, C)
    torch.cuda.synchronize()
    assert dg.last_repaired > 0
    fn, fc = oracle.triangulate_box(moved, box, 1.0, 16)
    got = T.cpu().numpy().reshape(n, dg.MaxSize)
    assert np.array_equal(canon(got, C.cpu().numpy()), canon(fn, fc))


def _local_canon(bd, tri, num, s):
    nb = tri.cpu().numpy().reshape(bd.N, bd.MaxSize)
    cnt = num.cpu().numpy()
    a, b = bd.offsets[s], bd.offsets[s + 1]
    cs = cnt[a:b]
    loc = np.where(np.arange(bd.MaxSize)[None, :] < cs[:, None], nb[a:b] - a, -1)
    return rings.canonical(loc, cs)


def mixed_batch():
    boxes = [(40.0, 40.0), BOXES["shear"], BOXES["general"], (30.0, 20.0), BOXES["lefthanded"]]
    sizes = [1600, 1998, 2000, 600, 2016]
    pts = []
    for s, (b, n) in enumerate(zip(boxes, sizes)):
        pts.append(inputs.uniform_points(n, b[0], b[1], seed=60 + s) if len(b) == 2 else
                   inputs.general_points(n, b, seed=60 + s))
    return boxes, sizes, pts


def test_mixed_batch_equals_each_set_alone():
    boxes, sizes, pts = mixed_batch()
    bd = BatchedDelaunayGPU(sizes, boxes, 16, 1.0)
    assert bd.general
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    T = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    C = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, T, C)
    torch.cuda.synchronize()
    for s, (b, xy) in enumerate(zip(boxes, pts)):
        mat = (b[0], 0.0, 0.0, b[1]) if len(b) == 2 else b
        on, oc = oracle.triangulate_box(xy, mat, 1.0, 16)
        assert np.array_equal(_local_canon(bd, T, C, s), canon(on, oc)), s


@pytest.mark.parametrize("use_graph", [True, False])
def test_brownian_loop_in_general_boxes(use_graph):
    boxes, sizes, pts = mixed_batch()
    K, sigma, seed = 10, 0.02, 77
    bd = BatchedDelaunayGPU(sizes, boxes, 16, 1.0)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    T = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    C = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, T, C)
    print(f"[general box] md graph={use_graph}", flush=True)
    rep = bd.md_run(X, T, C, K, sigma, seed, use_graph=use_graph, check_every=5)
    torch.cuda.synchronize()
    assert rep["steps"] == K and rep["flagged_total"] > 0
    got = X.cpu().numpy()
    for s, (b, xy) in enumerate(zip(boxes, pts)):
        mat = (b[0], 0.0, 0.0, b[1]) if len(b) == 2 else b
        for t in range(K):
            xy = (inputs.md_step(xy, b[0], b[1], sigma, seed, t, s) if len(b) == 2 else
                  inputs.md_step_general(xy, b, sigma, seed, t, s))
        a, e = bd.offsets[s], bd.offsets[s + 1]
        assert np.array_equal(got[a:e], xy), s
        on, oc = oracle.triangulate_box(xy, mat, 1.0, 16)
        assert np.array_equal(_local_canon(bd, T, C, s), canon(on, oc)), s


def test_general_voronoi_vertices_and_energy():
    box = BOXES["general"]
    n = 3000
    xy = inputs.general_points(n, box, seed=46)
    nb, c, dg, (P, T, C) = gpu_tri_box(xy, box)
    ctx = Context(0)
    L = _lib.load()
    ms = dg.MaxSize
    V = torch.zeros(n * ms * 2, dtype=torch.float64, device="cuda")
    b = (ctypes.c_double * 4)(*box)
    st = L.gdt_voronoi_vertices_general(ctx.handle, P.data_ptr(), n, b, T.data_ptr(),
                                        C.data_ptr(), ms, V.data_ptr(), None)
    _lib.check(st, "gdt_voronoi_vertices_general")
    torch.cuda.synchronize()
    got = V.cpu().numpy().reshape(n, ms, 2)
    X = np.array(box).reshape(2, 2)
    Xi = np.linalg.inv(X)
    for i in range(0, n, 5):
        k = c[i]
        d = xy[nb[i, :k]] - xy[i]
        w = d @ Xi.T
        d = (w - np.round(w)) @ X.T
        a, bb = d, np.roll(d, -1, axis=0)
        an, bn = (a * a).sum(1), (bb * bb).sum(1)
        den = 2.0 * (a[:, 0] * bb[:, 1] - a[:, 1] * bb[:, 0])
        ref = np.stack([(an * bb[:, 1] - bn * a[:, 1]) / den, (bn * a[:, 0] - an * bb[:, 0]) / den],
                       1)
        assert np.allclose(got[i, :k], ref, rtol=1e-10, atol=1e-10)
    area, perim, energy = voronoi_forward(ctx, P, T, C, box, a0=1.0, p0=3.8)
    grad = voronoi_adjoint(ctx, P, T, C, box, area, perim, a0=1.0, p0=3.8)
    torch.cuda.synchronize()
    E, A, Pm, G = oracle.voronoi_energy_box(xy, box, nb, c, a0=1.0, p0=3.8)
    assert np.allclose(area.cpu().numpy(), A, rtol=1e-10, atol=1e-10)
    assert np.allclose(perim.cpu().numpy(), Pm, rtol=1e-10, atol=1e-10)
    assert abs(float(energy.cpu()[0]) - E) <= 1e-9 * abs(E)
    assert np.allclose(grad.cpu().numpy(), G, rtol=1e-8, atol=1e-8)


def test_small_general_box_defers_to_the_wavefront_kernel():
    """A box of about 7 x 7 cells: per-point searches that the grid's search
    radius cannot certify are redone by the wavefront kernel; rings equal
    Qhull and the CPU mode's rows."""
    box = (7.0, 2.0, 0.5, 7.0)
    n = 50
    xy = inputs.general_points(n, box, seed=40)
    nb, c, dg, _ = gpu_tri_box(xy, box)
    qn, qc = periodic_rings_box(xy, box, margin=2.0)
    assert np.array_equal(canon(nb, c), canon(qn, qc))
    hn, hc = host_tri_box(xy, box, dg.MaxSize)
    assert same_rows(nb, c, hn, hc)
