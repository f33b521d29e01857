"""Autotuner: every timed configuration (threads per block, cell scale, tile
or row-major cell order, LDS tile edge, lanes per point) gives the same
rings, the fastest is applied to the context, and the tuned triangulation
equals the oracle's (65,536 uniform points, box 256); the Brownian-loop
tuner (tile edge, lanes per repaired point, block) gives the same positions
and rings for every configuration and leaves the caller's arrays alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd import autotune  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import BatchedDelaunayGPU  # noqa: E402

KNOBS = ("block", "tiled", "cell_scale", "tile_shift", "lanes")


def test_tuned_configuration_is_applied_and_exact():
    L = 256.0
    xy = inputs.uniform_points(65536, L, L, seed=17)
    pts = torch.from_numpy(xy).cuda()
    autotune.clear_cache()
    cfgs = [(256, 1, 1.0), (128, 1, 0.9, 4), (320, 1, 1.15, 5, 2), (512, 1, 0.8, 4, 2),
            (256, 0, 1.0), dict(block=256, tiled=2, cell_scale=1.0, tile_shift=4, lanes=1)]
    dg, best = autotune.tuned_delaunay(pts, (L, L), precision=PREC_FP32F, configs=cfgs, reps=2)
    assert len(best["table"]) == len(cfgs)
    c = dg.ctx.config()
    for k in KNOBS:
        if k in best:
            assert getattr(c, k) == pytest.approx(best[k]), k
    n = xy.shape[0]
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    nb, cnt = tri.cpu().numpy().reshape(n, dg.MaxSize), num.cpu().numpy()
    onb, oc = oracle.triangulate(xy, L, L, 1.0, 16)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(onb, oc) == best["ring_hash"]
    # a second call hits the cache and re-applies the same configuration
    again = autotune.tune(dg, pts)
    assert again is best


def test_brownian_loop_tuner():
    sizes = [1024, 1500, 2048, 3000, 4096, 1100]
    pts = [inputs.uniform_points(n, float(np.sqrt(n)), seed=50 + s) for s, n in enumerate(sizes)]
    boxes = [float(np.sqrt(n)) for n in sizes]
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 16, 1.0)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)
    torch.cuda.synchronize()
    X0, t0, n0, ms0 = X.clone(), tri.clone(), num.clone(), bd.MaxSize
    autotune.clear_cache()
    cfgs = autotune.md_candidates(blocks=(256, 512))
    best = autotune.tune_md(bd, X, tri, num, configs=cfgs, steps=8, seed=3)
    assert len(best["table"]) == len(cfgs) == 12      # block x tile edge x lanes (1, 2, 4)
    c = bd.ctx.config()
    assert (c.block, c.tile_shift, c.repair_lanes) == (best["block"], best["tile_shift"],
                                                       best["repair_lanes"])
    assert torch.equal(X, X0) and torch.equal(tri, t0) and torch.equal(num, n0)
    assert bd.MaxSize == ms0
