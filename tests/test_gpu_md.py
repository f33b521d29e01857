"""Brownian test-and-repair loop on the device (gdt_md_run), eager and as
hipGraph replays:

* positions after K steps are bit-identical to the numpy restatement of the
  counter-based noise (inputs.md_step), and the rings equal the oracle's
  triangulation of those positions (canonical, per set);
* the graph run (K replays of one captured step, with readbacks between
  checks) equals the eager run for K = 1, 3 and 12;
* on the 256-set workload (about 3.92 M points, sigma = 0.01), 100 replayed
  steps end with rings equal to a fresh global triangulation of the final
  positions, and the full 1,000-step run ends with every set's rings equal to
  the oracle's triangulation of its final positions (the positions of the
  four smallest sets equal to the numpy restatement);
* a step whose repaired ring outgrows the row stride halts the loop on the
  device: no later step moves a point or touches a row, and the host's
  recovery (global triangulation of the positions, wider rows) continues
  from the next step, so the run still equals the restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import BatchedDelaunayGPU  # noqa: E402

SIGMA, SEED = 0.01, 2024


def canonical_sets(bd, tri, num):
    nb = tri.cpu().numpy().reshape(bd.N, bd.MaxSize)
    cnt = num.cpu().numpy()
    out = []
    for s in range(bd.nsets):
        a, b = bd.offsets[s], bd.offsets[s + 1]
        cs = cnt[a:b]
        loc = np.where(np.arange(bd.MaxSize)[None, :] < cs[:, None], nb[a:b] - a, -1)
        out.append((cs.copy(), rings.canonical(loc, cs)))
    return out


def same_sets(x, y):
    for (c1, k1), (c2, k2) in zip(x, y):
        if not (np.array_equal(c1, c2) and np.array_equal(k1, k2)):
            return False
    return len(x) == len(y)


def small_batch():
    sizes = [1024, 1500, 2048, 3000, 4096, 1100]
    pts = [inputs.uniform_points(n, float(np.sqrt(n)), seed=50 + s) for s, n in enumerate(sizes)]
    boxes = [float(np.sqrt(n)) for n in sizes]
    return sizes, pts, boxes


def start(sizes, pts, boxes, prec=PREC_FP64):
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 16, 1.0, precision=prec)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)
    torch.cuda.synchronize()
    return bd, X, tri, num


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_positions_and_rings_match_restatement(prec):
    sizes, pts, boxes = small_batch()
    K = 12
    bd, X, tri, num = start(sizes, pts, boxes, prec)
    rep = bd.md_run(X, tri, num, K, SIGMA, SEED, use_graph=True, check_every=4)
    torch.cuda.synchronize()
    assert rep["steps"] == K and rep["recoveries"] == 0 and rep["flagged_total"] > 0
    ref_pos, ref_rings = [], []
    for s, (xy, L) in enumerate(zip(pts, boxes)):
        for t in range(K):
            xy = inputs.md_step(xy, L, L, SIGMA, SEED, t, s)
        ref_pos.append(xy)
        nb, c = oracle.triangulate(xy, L, L, 1.0, 20)
        ref_rings.append((c, rings.canonical(nb, c)))
    got = X.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), np.concatenate(ref_pos).view(np.uint64))
    assert same_sets(canonical_sets(bd, tri, num), ref_rings)


@pytest.mark.parametrize("K", [1, 3, 12])
def test_graph_replays_equal_eager(K):
    sizes, pts, boxes = small_batch()
    bd_g, Xg, tg, ng = start(sizes, pts, boxes)
    bd_e, Xe, te, ne = start(sizes, pts, boxes)
    rg = bd_g.md_run(Xg, tg, ng, K, SIGMA, SEED, use_graph=True, check_every=2)
    re = bd_e.md_run(Xe, te, ne, K, SIGMA, SEED, use_graph=False, check_every=2)
    torch.cuda.synchronize()
    assert rg == re
    assert torch.equal(Xg, Xe)
    assert torch.equal(ng, ne)
    assert same_sets(canonical_sets(bd_g, tg, ng), canonical_sets(bd_e, te, ne))


def test_c4_hundred_steps_equal_fresh_triangulation():
    sizes = inputs.batch_sizes(256)
    pts, boxes = inputs.batch_points(sizes)
    bd, X, tri, num = start(sizes, pts, boxes)
    rep = bd.md_run(X, tri, num, 100, SIGMA, SEED, use_graph=True, check_every=50)
    torch.cuda.synchronize()
    assert rep["steps"] == 100
    assert 0 < rep["flagged_total"] < 100 * bd.N
    fresh = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 16, 1.0)
    tri2 = torch.zeros(bd.N * fresh.MaxSize, dtype=torch.int32, device="cuda")
    num2 = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    fresh.globalDelaunayTriangulation(X, tri2, num2)
    torch.cuda.synchronize()
    assert int(num.sum().item()) == 6 * bd.N
    assert same_sets(canonical_sets(bd, tri, num), canonical_sets(fresh, tri2, num2))


def test_overflow_inside_the_loop_is_recovered_exactly():
    """Rows sized to the starting triangulation (guess 4, grown only to the
    initial longest ring, no headroom) overflow once large moves lengthen
    some rings; the device halts the loop at the end of the failing step
    (the positions are exactly that step's, no later step ran), the host
    grows MaxSize, re-triangulates globally and goes on; the final positions
    and rings still equal the restatement."""
    sizes, pts, boxes = small_batch()
    sigma, K = 0.25, 12
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 4, 1.0)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)
    ms0 = bd.MaxSize
    rep = bd.md_run(X, tri, num, K, sigma, SEED, use_graph=True, check_every=1, headroom=0)
    torch.cuda.synchronize()
    assert rep["steps"] == K
    assert rep["recoveries"] >= 1 and bd.MaxSize > ms0
    assert all(r["status"] == "ring longer than the row stride" for r in rep["recovered"])
    assert rep["skipped_replays"] == 0      # checked after every step: nothing to skip
    ref_pos, ref_rings = [], []
    for s, (xy, L) in enumerate(zip(pts, boxes)):
        for t in range(K):
            xy = inputs.md_step(xy, L, L, sigma, SEED, t, s)
        ref_pos.append(xy)
        nb, c = oracle.triangulate(xy, L, L, 1.0, 24)
        ref_rings.append((c, rings.canonical(nb, c)))
    got = X.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), np.concatenate(ref_pos).view(np.uint64))
    assert same_sets(canonical_sets(bd, tri, num), ref_rings)


def test_halted_loop_runs_no_step_after_the_failure():
    """With checks only every 6 steps, the replays after an overflow inside
    a check interval are enqueued but must change nothing: the positions
    returned with the failure are exactly those of the failing step (the
    restatement's, step for step) and the recovery continues from there."""
    sizes, pts, boxes = small_batch()
    sigma, K = 0.25, 12
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 4, 1.0)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)
    rep = bd.md_run(X, tri, num, K, sigma, SEED, use_graph=True, check_every=6, headroom=0)
    torch.cuda.synchronize()
    assert rep["steps"] == K and rep["recoveries"] >= 1
    assert rep["skipped_replays"] >= 1, "no replay followed a failure inside a check interval"
    ref_pos = []
    for s, (xy, L) in enumerate(zip(pts, boxes)):
        for t in range(K):
            xy = inputs.md_step(xy, L, L, sigma, SEED, t, s)
        ref_pos.append(xy)
    got = X.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), np.concatenate(ref_pos).view(np.uint64))
    ref = []
    for xy, L in zip(ref_pos, boxes):
        nb, c = oracle.triangulate(xy, L, L, 1.0, 24)
        ref.append((c, rings.canonical(nb, c)))
    assert same_sets(canonical_sets(bd, tri, num), ref)


def test_c4_thousand_steps_equal_oracle():
    """The full C4 run: 256 sets, 1,000 hipGraph-replayed steps of sigma =
    0.01 from the batched triangulation, rows widened to the longest ring + 4
    before the loop.  Every set's final rings equal the oracle's
    triangulation of its final positions, and the four smallest sets'
    positions equal the numpy restatement of the noise, step for step."""
    sizes = inputs.batch_sizes(256)
    pts, boxes = inputs.batch_points(sizes)
    bd, X, tri, num = start(sizes, pts, boxes)
    K = 1000
    rep = bd.md_run(X, tri, num, K, SIGMA, SEED, use_graph=True, check_every=100)
    torch.cuda.synchronize()
    assert rep["steps"] == K
    got = X.cpu().numpy()
    canon = canonical_sets(bd, tri, num)
    for s in np.argsort(sizes, kind="stable")[:4]:
        xy, L = pts[s], boxes[s]
        for t in range(K):
            xy = inputs.md_step(xy, L, L, SIGMA, SEED, t, int(s))
        a, b = bd.offsets[s], bd.offsets[s + 1]
        assert np.array_equal(got[a:b].view(np.uint64), xy.view(np.uint64)), f"set {s}"
    for s in range(bd.nsets):
        a, b = bd.offsets[s], bd.offsets[s + 1]
        L = boxes[s]
        nb, c = oracle.triangulate(np.ascontiguousarray(got[a:b]), L, L, 1.0, 24)
        assert np.array_equal(canon[s][0], c), f"set {s}: counts differ from the oracle"
        assert np.array_equal(canon[s][1], rings.canonical(nb, c)), f"set {s}: rings differ"
