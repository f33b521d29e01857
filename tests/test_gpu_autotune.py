"""Autotuner: every timed configuration (threads per block, cell scale, tile
or row-major cell order) gives the same rings, the fastest is applied to the
context, and the tuned triangulation equals the oracle's (65,536 uniform
points, box 256)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd import autotune  # noqa: E402


def test_tuned_configuration_is_applied_and_exact():
    L = 256.0
    xy = inputs.uniform_points(65536, L, L, seed=17)
    pts = torch.from_numpy(xy).cuda()
    autotune.clear_cache()
    cfgs = [(256, 1, 1.0), (128, 1, 0.9), (320, 1, 1.15), (512, 1, 0.8), (256, 0, 1.0)]
    dg, best = autotune.tuned_delaunay(pts, (L, L), precision=PREC_FP32F, configs=cfgs, reps=2)
    assert len(best["table"]) == len(cfgs)
    c = dg.ctx.config()
    assert (c.block, c.tiled, c.cell_scale) == (best["block"], best["tiled"], best["cell_scale"])
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
