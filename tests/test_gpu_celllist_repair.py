"""GPU cell list (reference layout export), triangle enumeration, circumcircle
test and local repair, against the oracle."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpu_util import gpu_triangulate  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, _lib, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import CellListGPU, PeriodicBoundaries  # noqa: E402


@pytest.mark.parametrize("n,L,cell,clump", [(4096, 64.0, 1.0, 0), (4096, 64.0, 0.25, 0),
                                           (8192, 64.0, 1.0, 2000), (16384, 128.0, 1.0, 3000),
                                           (65536, 256.0, 1.0, 0),
                                           (65536, 256.0, 1.0, 5000)])
def test_cell_list_export_matches_reference_binning(n, L, cell, clump):
    """Up to 16,384 points the build is one launch (one block per 1,024 keys),
    above it four passes; cell 0.25 gives the one-launch build its most
    blocks (64), a clump puts hundreds of points in a cell."""
    print(f"[cells] n {n} box {L} cell {cell} clump {clump}", flush=True)
    xy = inputs.uniform_points(n, L, L, seed=1)
    if clump:
        xy[:clump] = 30.0 + 3.0 * inputs.uniform_points(clump, 1.0, 1.0, seed=2)
    cl = CellListGPU(PeriodicBoundaries(L, L))
    cl.setGridSize(cell)
    cl.computeGPU(torch.from_numpy(xy).cuda())
    sizes = cl.cell_sizes.cpu().numpy()
    idxs = cl.idxs.cpu().numpy().reshape(sizes.size, cl.getNmax())
    # reference binning in index order (cellListGPU.cpp:306-326)
    m = int(L / cell)
    bx = np.floor(xy[:, 0] / (L / m)).astype(int)
    by = np.floor(xy[:, 1] / (L / m)).astype(int)
    key = by * m + bx
    assert cl.getXsize() == m and cl.getYsize() == m
    assert np.array_equal(sizes, np.bincount(key, minlength=m * m))
    assert cl.getNmax() == sizes.max()
    order = np.argsort(key, kind="stable")
    start = np.concatenate([[0], np.cumsum(sizes)])
    for c in np.nonzero(sizes)[0]:
        assert idxs[c, :sizes[c]].tolist() == order[start[c]:start[c + 1]].tolist()


def _triangles(dg, nb, c):
    L = _lib.load()
    n, stride = nb.shape
    tn = torch.from_numpy(np.ascontiguousarray(nb)).cuda()
    tc = torch.from_numpy(np.ascontiguousarray(c)).cuda()
    cap = 2 * n + 64
    tris = torch.zeros(3 * cap, dtype=torch.int32, device="cuda")
    nt, tot = ctypes.c_int(), ctypes.c_longlong()
    _lib.check(L.gdt_triangles(dg.ctx.handle, tn.data_ptr(), tc.data_ptr(), n, stride,
                               tris.data_ptr(), cap, ctypes.byref(nt), ctypes.byref(tot),
                               torch.cuda.current_stream().cuda_stream), "gdt_triangles")
    return tris[:3 * nt.value].cpu().numpy().reshape(-1, 3), tot.value


def test_triangles_match_oracle():
    xy = inputs.uniform_points(4096, 64.0, 64.0, seed=1)
    nb, c, dg = gpu_triangulate(xy, 64.0, 64.0, PREC_FP64)
    tris, tot = _triangles(dg, nb, c)
    assert tris.shape[0] == 2 * 4096 and tot == 6 * 4096
    on, oc = oracle.triangulate(xy, 64.0, 64.0, 1.0, 16)
    otris, _ = oracle.triangles(on, oc)

    def canon(t):
        t = t.copy()
        for r in t:          # keep CCW order, rotate the smallest index first
            k = int(np.argmin(r))
            r[:] = np.roll(r, -k)
        return t[np.lexsort(t.T[::-1])]
    assert np.array_equal(canon(tris), canon(otris))
    assert np.all(tris[:, 0] < tris[:, 1]) and np.all(tris[:, 0] < tris[:, 2])


def _moved(xy, L, frac, seed, amp=0.1):
    rng = np.random.default_rng(seed)
    n = xy.shape[0]
    sel = rng.choice(n, size=int(frac * n), replace=False)
    m = xy.copy()
    m[sel] += rng.uniform(-amp, amp, size=(sel.size, 2))
    return inputs.wrap_into_box(m, L, L)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_test_and_repair_equals_fresh_triangulation(prec):
    """triangulation.cpp:349-413 (-z -2): move 5% of the points by U(-0.1,0.1),
    test-and-repair, compare with a fresh triangulation of the moved points."""
    L = 128.0
    xy = inputs.uniform_points(16384, L, L, seed=31)
    nb, c, dg = gpu_triangulate(xy, L, L, prec)
    moved = _moved(xy, L, 0.05, 32)
    pts = torch.from_numpy(moved).cuda()
    tri = torch.from_numpy(nb.reshape(-1).copy()).cuda()
    num = torch.from_numpy(c.copy()).cuda()
    dg.testAndRepairDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    assert dg.last_repaired > 0
    rn = tri.cpu().numpy().reshape(16384, dg.MaxSize)
    rc = num.cpu().numpy()
    fn, fc = oracle.triangulate(moved, L, L, 1.0, 16)
    assert rings.ring_hash(rn, rc) == rings.ring_hash(fn, fc)
    # and the oracle's own test-and-repair agrees
    on, oc = oracle.triangulate(xy, L, L, 1.0, 16)
    tn, tc, _ = oracle.test_and_repair(moved, L, L, on, oc)
    assert rings.ring_hash(tn, tc) == rings.ring_hash(fn, fc)


def test_valid_triangulation_flags_nothing():
    L = 64.0
    xy = inputs.uniform_points(4096, L, L, seed=1)
    nb, c, dg = gpu_triangulate(xy, L, L, PREC_FP64)
    tri = torch.from_numpy(nb.reshape(-1).copy()).cuda()
    num = torch.from_numpy(c.copy()).cuda()
    dg.testAndRepairDelaunayTriangulation(torch.from_numpy(xy).cuda(), tri, num)
    assert dg.last_repaired == 0
    assert np.array_equal(tri.cpu().numpy().reshape(nb.shape), nb)


def test_circumcircle_test_right_after_global_triangulation():
    """The global triangulation builds its cell list without the caller ->
    slot map; the circumcircle test on that cell list (no rebuild) fills the
    map first and flags nothing for the valid triangulation."""
    L = 64.0
    xy = inputs.uniform_points(4096, L, L, seed=4)
    nb, c, dg = gpu_triangulate(xy, L, L, PREC_FP64)
    tris, _ = _triangles(dg, nb, c)
    pts = torch.from_numpy(xy).cuda()
    tt = torch.from_numpy(tris.reshape(-1).copy()).cuda()
    rep = torch.zeros(4096, dtype=torch.int32, device="cuda")
    nf = ctypes.c_int(-1)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().gdt_test_circumcircles(dg.ctx.handle, pts.data_ptr(), 4096,
                                                  tt.data_ptr(), tris.shape[0], rep.data_ptr(),
                                                  ctypes.byref(nf), s), "test")
    assert nf.value == 0
    assert (rep.cpu().numpy() < 0).all()


def _bad_triangles(xy, L, tris):
    """Triangles whose circumcircle strictly contains another point (scipy
    periodic KD-tree; independent of both the oracle and the kernels)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(xy, boxsize=[L, L])
    v = xy[tris[:, 0]]
    a = xy[tris[:, 1]] - v
    b = xy[tris[:, 2]] - v
    a -= L * np.round(a / L)
    b -= L * np.round(b / L)
    d = 2 * (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    aa, bb = (a ** 2).sum(1), (b ** 2).sum(1)
    q = np.stack([(aa * b[:, 1] - bb * a[:, 1]) / d, (bb * a[:, 0] - aa * b[:, 0]) / d], 1)
    R = np.sqrt((q ** 2).sum(1))
    centres = np.mod(v + q, L)
    inside = tree.query_ball_point(centres, R * (1 - 1e-9))
    bad = np.array([len(set(p) - set(t.tolist())) > 0 for p, t in zip(inside, tris)])
    near = tree.query_ball_point(centres, R * (1 + 1e-7))
    return bad, near


def test_circumcircle_known_answer():
    """Move one point to the centroid of a neighbouring triangle: every
    triangle whose circle now holds a point has its three vertices flagged
    (DelaunayGPU.cu:277-282), and nothing else is flagged except points
    inside (or on) such circles.  Which intruder gets flagged depends on the
    scan order, so intruders are not compared with the oracle's."""
    L = 64.0
    xy = inputs.uniform_points(4096, L, L, seed=1)
    nb, c, dg = gpu_triangulate(xy, L, L, PREC_FP64)
    tris, _ = _triangles(dg, nb, c)
    i, a, b = tris[100]
    k = [p for p in nb[a, :c[a]] if p not in (i, a, b)][0]
    moved = xy.copy()
    moved[k] = (xy[i] + xy[a] + xy[b]) / 3.0
    pts = torch.from_numpy(moved).cuda()
    Lb = _lib.load()
    h = dg.ctx.handle
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(Lb.gdt_celllist_build(h, pts.data_ptr(), 4096, L, L, 1.0, s), "celllist")
    tt = torch.from_numpy(tris.reshape(-1).copy()).cuda()
    rep = torch.zeros(4096, dtype=torch.int32, device="cuda")
    nf = ctypes.c_int()
    _lib.check(Lb.gdt_test_circumcircles(h, pts.data_ptr(), 4096, tt.data_ptr(), tris.shape[0],
                                         rep.data_ptr(), ctypes.byref(nf), s), "test")
    r = rep.cpu().numpy()
    flagged = set(np.nonzero(r >= 0)[0].tolist())
    assert all(r[v] == v for v in (i, a, b, k))
    bad, near = _bad_triangles(moved, L, tris)
    must = set(tris[bad].reshape(-1).tolist())
    may = must | {p for j in np.nonzero(bad)[0] for p in near[j]}
    assert must <= flagged <= may
    assert nf.value == len(flagged)
    # the oracle flags the same vertex set (its intruders may differ)
    want = oracle.test_circumcircles(moved, L, L, tris)
    assert must <= set(np.nonzero(want >= 0)[0].tolist()) <= may


def test_tiled_test_flags_within_known_answer():
    """The per-tile LDS test inside testAndRepair (triangles in slot order)
    flags between the vertices of the truly bad triangles and those plus the
    points inside or on their circles, and the repaired rings equal the
    oracle's triangulation of the moved points."""
    L = 64.0
    xy = inputs.uniform_points(4096, L, L, seed=1)
    nb, c, dg = gpu_triangulate(xy, L, L, PREC_FP64)
    tris, _ = _triangles(dg, nb, c)
    i, a, b = tris[100]
    k = [p for p in nb[a, :c[a]] if p not in (i, a, b)][0]
    moved = xy.copy()
    moved[k] = (xy[i] + xy[a] + xy[b]) / 3.0
    bad, near = _bad_triangles(moved, L, tris)
    must = set(tris[bad].reshape(-1).tolist())
    may = must | {p for j in np.nonzero(bad)[0] for p in near[j]}
    tri = torch.from_numpy(nb.reshape(-1).copy()).cuda()
    num = torch.from_numpy(c.copy()).cuda()
    dg.testAndRepairDelaunayTriangulation(torch.from_numpy(moved).cuda(), tri, num)
    assert len(must) <= dg.last_repaired <= len(may)
    on, oc = oracle.triangulate(moved, L, L, 1.0, 16)
    got = tri.cpu().numpy().reshape(4096, dg.MaxSize)
    assert rings.ring_hash(got, num.cpu().numpy()) == rings.ring_hash(on, oc)


def test_euler_check_reports_corrupted_rings():
    L = 64.0
    xy = inputs.uniform_points(4096, L, L, seed=1)
    nb, c, dg = gpu_triangulate(xy, L, L, PREC_FP64)
    c2 = c.copy()
    c2[7] -= 1
    tri = torch.from_numpy(nb.reshape(-1).copy()).cuda()
    num = torch.from_numpy(c2).cuda()
    with pytest.raises(_lib.GdtError) as e:
        dg.testAndRepairDelaunayTriangulation(torch.from_numpy(xy).cuda(), tri, num)
    assert e.value.status == _lib.GDT_ERR_EULER
