"""Exact ties on the GPU (C ABI via the DelaunayGPU mirror): exact square
lattices, where every square's four points are cocircular and every point has
four equidistant nearest neighbours.  The rings must be the Delaunay
triangulation of the symbolically perturbed points (lifted coordinate
|p|^2 + eps^(1 + caller index)), the same for both precisions and cell
orders, and the circumcircle test-and-repair must leave them unchanged.  The
reference defines no tie rule, so these cases are checked against exact
arithmetic, not against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from exact_delaunay import (SCALE, perturbed_violations, ring_inconsistencies,  # noqa: E402
                            square_lattice)
from gpu_util import gpu_triangulate  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, rings  # noqa: E402

# (nx, ny, spacing in cells)
LATTICES = {"square64": (64, 64, 1.0), "rect48x80": (48, 80, 1.0), "half64": (64, 64, 0.5),
            "quarter96": (96, 96, 0.25)}


@pytest.mark.parametrize("case", sorted(LATTICES))
def test_lattice_rings_are_the_perturbed_delaunay_triangulation(case):
    nx, ny, sp = LATTICES[case]
    P = square_lattice(nx, ny, sp)
    xy, lx, ly = P / SCALE, nx * sp, ny * sp
    nb, c, dg = gpu_triangulate(xy, lx, ly, PREC_FP64)
    assert int(c.sum()) == 6 * len(c)
    assert ring_inconsistencies(nb, c) == 0
    assert perturbed_violations(P, lx, ly, nb, c, D=3.0 * sp) == 0
    ref = rings.ring_hash(nb, c)
    n32, c32, _ = gpu_triangulate(xy, lx, ly, PREC_FP32F, dg=dg)
    assert rings.ring_hash(n32, c32) == ref
    for tiled, block in ((0, 128), (2, 512)):
        dg.ctx.set_config(tiled=tiled, block=block)
        for prec in (PREC_FP64, PREC_FP32F):
            nt, ct, _ = gpu_triangulate(xy, lx, ly, prec, dg=dg)
            assert rings.ring_hash(nt, ct) == ref, (tiled, prec)


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_lattice_test_and_repair_keeps_the_rings(prec):
    """Cocircular triangles are flagged by the conservative circumcircle test
    and recomputed: the repaired rings equal the global ones."""
    nx, ny, sp = LATTICES["square64"]
    P = square_lattice(nx, ny, sp)
    xy, lx, ly = P / SCALE, nx * sp, ny * sp
    nb, c, dg = gpu_triangulate(xy, lx, ly, prec)
    tri = torch.from_numpy(nb.reshape(-1).copy()).cuda()
    num = torch.from_numpy(c.copy()).cuda()
    dg.testAndRepairDelaunayTriangulation(torch.from_numpy(xy).cuda(), tri, num)
    torch.cuda.synchronize()
    rn = tri.cpu().numpy().reshape(nb.shape)
    rc = num.cpu().numpy()
    assert rings.ring_hash(rn, rc) == rings.ring_hash(nb, c)
