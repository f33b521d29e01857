"""The tuner's knobs change the work split, never the rings: LDS tiles of
16x16 cells (tile_shift 4) or 32x32 (5), one, two or four lanes per point in the
global triangulation (lanes; 64: the star kernel) and in repairs
(repair_lanes), each against the
oracle: uniform points (the 2^20 workload against its golden hash), a
rectangular box, a cluster that overflows the 16x16 windows, test-and-repair
and a short Brownian run."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import (BatchedDelaunayGPU, DelaunayGPU,  # noqa: E402
                                                   PeriodicBoundaries)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KNOBS = [(4, 1), (4, 2), (5, 2), (4, 4), (4, 64)]  # (tile_shift, lanes); (5, 1) is the default


def triangulate(xy, lx, ly, prec, ts, lanes, repair_lanes=2):
    n = xy.shape[0]
    dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(lx, ly))
    dg.setPrecision(prec)
    dg.ctx.set_config(tile_shift=ts, lanes=lanes, repair_lanes=repair_lanes)
    pts = torch.from_numpy(np.ascontiguousarray(xy)).cuda()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    return dg, pts, tri, num


def as_np(dg, tri, num):
    n = num.numel()
    return tri.cpu().numpy().reshape(n, dg.MaxSize), num.cpu().numpy()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
@pytest.mark.parametrize("ts,lanes", KNOBS)
def test_knobs_uniform_and_rectangular(ts, lanes, prec):
    for n, lx, ly, seed in ((65536, 256.0, 256.0, 61), (20000, 200.0, 100.0, 62), (4096, 64.0, 64.0, 1)):
        print(f"[knobs] ts {ts} lanes {lanes} prec {prec}: n {n} box {lx}x{ly}", flush=True)
        xy = inputs.uniform_points(n, lx, ly, seed=seed)
        dg, _, tri, num = triangulate(xy, lx, ly, prec, ts, lanes)
        c = dg.ctx.config()
        assert (c.tile_shift, c.lanes) == (ts, lanes)
        on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
        assert rings.ring_hash(*as_np(dg, tri, num)) == rings.ring_hash(on, oc)


@pytest.mark.parametrize("ts,lanes", KNOBS)
def test_knobs_c2_golden(ts, lanes):
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        want = json.load(fh)["workloads"]["c2"]["ring_sha256"]
    xy = inputs.uniform_points(1 << 20, 1024.0, 1024.0, seed=2)
    print(f"[knobs] c2 ts {ts} lanes {lanes}", flush=True)
    dg, _, tri, num = triangulate(xy, 1024.0, 1024.0, PREC_FP32F, ts, lanes)
    assert rings.ring_hash(*as_np(dg, tri, num)) == want


def test_knobs_cluster_overflows_small_windows():
    """3,000 points in a 6x6 square overflow the 1,024-point windows of the
    16x16 tiles around it: those tiles defer to the wavefront kernel."""
    L = 141.0
    base = inputs.uniform_points(17000, L, L, seed=41)
    cluster = 60.0 + np.random.default_rng(42).uniform(0.0, 6.0, size=(3000, 2))
    xy = np.concatenate([base, cluster])
    on, oc = oracle.triangulate(xy, L, L, 1.0, 16)
    for lanes in (1, 2, 4):
        print(f"[knobs] cluster ts 4 lanes {lanes}", flush=True)
        dg, _, tri, num = triangulate(xy, L, L, PREC_FP64, 4, lanes)
        assert dg.stats()["deferred"] > 0
        assert rings.ring_hash(*as_np(dg, tri, num)) == rings.ring_hash(on, oc)


@pytest.mark.parametrize("ts,repair_lanes", [(4, 1), (4, 2), (5, 1), (5, 4)])
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_knobs_test_and_repair(ts, repair_lanes, prec):
    n, L = 16384, 128.0
    xy = inputs.uniform_points(n, L, L, seed=63)
    print(f"[knobs] repair ts {ts} repair_lanes {repair_lanes} prec {prec}", flush=True)
    dg, pts, tri, num = triangulate(xy, L, L, prec, ts, 1, repair_lanes)
    moved = inputs.wrap_into_box(xy + 0.05 * inputs.gaussian(64, 2 * n).reshape(n, 2), L, L)
    dg.testAndRepairDelaunayTriangulation(torch.from_numpy(moved).cuda(), tri, num)
    torch.cuda.synchronize()
    assert dg.last_repaired > 0
    on, oc = oracle.triangulate(moved, L, L, 1.0, 16)
    assert rings.ring_hash(*as_np(dg, tri, num)) == rings.ring_hash(on, oc)


def test_knobs_brownian_loop():
    """12 replayed steps with 16x16 tiles and one lane per repaired point:
    positions equal the numpy restatement, rings the oracle's."""
    sizes = [1024, 1500, 2048, 3000, 4096, 1100]
    pts = [inputs.uniform_points(n, float(np.sqrt(n)), seed=50 + s) for s, n in enumerate(sizes)]
    boxes = [float(np.sqrt(n)) for n in sizes]
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 16, 1.0)
    bd.ctx.set_config(tile_shift=4, repair_lanes=1)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)
    print("[knobs] brownian ts 4", flush=True)
    K, sigma, seed = 12, 0.01, 2024
    rep = bd.md_run(X, tri, num, K, sigma, seed, use_graph=True, check_every=4)
    torch.cuda.synchronize()
    assert rep["steps"] == K and rep["flagged_total"] > 0
    got = X.cpu().numpy()
    nb = tri.cpu().numpy().reshape(bd.N, bd.MaxSize)
    cnt = num.cpu().numpy()
    for s, (xy, L) in enumerate(zip(pts, boxes)):
        for t in range(K):
            xy = inputs.md_step(xy, L, L, sigma, seed, t, s)
        a, b = bd.offsets[s], bd.offsets[s + 1]
        assert np.array_equal(got[a:b].view(np.uint64), xy.view(np.uint64))
        on, oc = oracle.triangulate(xy, L, L, 1.0, 24)
        loc = np.where(np.arange(bd.MaxSize)[None, :] < cnt[a:b, None], nb[a:b] - a, -1)
        assert np.array_equal(cnt[a:b], oc)
        assert np.array_equal(rings.canonical(loc, cnt[a:b]), rings.canonical(on, oc))
