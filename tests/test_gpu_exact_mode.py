"""Exact precision mode on the GPU (GDT_PREC_EXACT = 128): the per-point
kernels with every decision exact (gdt_exact_ring.h).

* near-degenerate lattices (non-dyadic spacing; exact lattices moved by a few
  ulps): rings bit-identical to the CPU mode's exact rings, consistent,
  summing to 6N, and exactly Delaunay with ties perturbed
  (tests/exact_rational.py, exact integer arithmetic);
* generic input: rings equal the oracle's (the FP64 path's);
* test-and-repair after moving points by a few ulps equals a fresh exact
  triangulation and is exactly Delaunay;
* a Brownian run of a small batch in exact mode: positions equal the numpy
  restatement, rings equal the CPU mode's exact rings of the final positions;
* a general box is refused.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_EXACT, _lib, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import (BatchedDelaunayGPU, DelaunayGPU,  # noqa: E402
                                                   PeriodicBoundaries)

from exact_delaunay import ring_inconsistencies  # noqa: E402
from exact_rational import exact_violations  # noqa: E402
from test_exact_mode import NEAR, host_tri, lattice  # noqa: E402


def gpu_tri(xy, lx, ly, cell, stride=24):
    n = xy.shape[0]
    dg = DelaunayGPU(n, stride, cell, PeriodicBoundaries(lx, ly))
    dg.setPrecision(PREC_EXACT)
    pts = torch.from_numpy(np.ascontiguousarray(xy)).cuda()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    return dg, pts, tri, num


def as_np(dg, tri, num):
    n = num.numel()
    return tri.cpu().numpy().reshape(n, dg.MaxSize), num.cpu().numpy()


@pytest.mark.parametrize("name", sorted(NEAR))
def test_gpu_exact_mode_near_degenerate(name):
    (xy, lx, ly), cell = NEAR[name]
    print(f"[exact] {name}: n {xy.shape[0]}", flush=True)
    dg, _, tri, num = gpu_tri(xy, lx, ly, cell)
    nb, cnt = as_np(dg, tri, num)
    assert int(cnt.sum()) == 6 * xy.shape[0]
    assert ring_inconsistencies(nb, cnt) == 0
    assert exact_violations(xy, lx, ly, nb, cnt, 3.0 * cell) == 0
    st, hnb, hcnt = host_tri(xy, lx, ly, PREC_EXACT, cell)
    assert st == 0
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(hnb, hcnt)


def test_gpu_exact_mode_generic_equals_oracle():
    for n, lx, ly, seed in ((65536, 256.0, 256.0, 61), (20000, 200.0, 100.0, 62)):
        print(f"[exact] generic n {n}", flush=True)
        xy = inputs.uniform_points(n, lx, ly, seed=seed)
        dg, _, tri, num = gpu_tri(xy, lx, ly, 1.0)
        on, oc = oracle.triangulate(xy, lx, ly, 1.0, 24)
        assert rings.ring_hash(*as_np(dg, tri, num)) == rings.ring_hash(on, oc)


def test_gpu_exact_mode_test_and_repair():
    xy, lx, ly = lattice(20, 20, 0.7)
    print("[exact] test-and-repair, lattice 0.7", flush=True)
    dg, pts, tri, num = gpu_tri(xy, lx, ly, 0.7)
    rng = np.random.default_rng(5)
    moved = xy.copy()
    pick = rng.choice(xy.shape[0], 60, replace=False)
    moved[pick] += rng.integers(-4, 5, size=(60, 2)) * np.spacing(moved[pick])
    dg.testAndRepairDelaunayTriangulation(torch.from_numpy(moved).cuda(), tri, num)
    torch.cuda.synchronize()
    nb, cnt = as_np(dg, tri, num)
    assert ring_inconsistencies(nb, cnt) == 0
    assert exact_violations(moved, lx, ly, nb, cnt, 3.0 * 0.7) == 0
    _, _, tri2, num2 = gpu_tri(moved, lx, ly, 0.7, stride=dg.MaxSize)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(*as_np(dg, tri2, num2))


def test_gpu_exact_mode_brownian_batch():
    sizes = [400, 625, 900]
    pts = [inputs.uniform_points(n, float(np.sqrt(n)), seed=70 + s) for s, n in enumerate(sizes)]
    boxes = [float(np.sqrt(n)) for n in sizes]
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 16, 1.0, precision=PREC_EXACT)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)
    print("[exact] brownian batch", flush=True)
    K, sigma, seed = 6, 0.01, 77
    rep = bd.md_run(X, tri, num, K, sigma, seed, use_graph=True, check_every=3)
    torch.cuda.synchronize()
    assert rep["steps"] == K
    got = X.cpu().numpy()
    nb = tri.cpu().numpy().reshape(bd.N, bd.MaxSize)
    cnt = num.cpu().numpy()
    for s, (xy, L) in enumerate(zip(pts, boxes)):
        for t in range(K):
            xy = inputs.md_step(xy, L, L, sigma, seed, t, s)
        a, b = bd.offsets[s], bd.offsets[s + 1]
        assert np.array_equal(got[a:b], xy)
        st, hnb, hcnt = host_tri(xy, L, L, PREC_EXACT, stride=bd.MaxSize)
        assert st == 0
        local = nb[a:b, :].copy()
        local[local >= 0] -= a
        assert rings.ring_hash(local, cnt[a:b]) == rings.ring_hash(hnb, hcnt)


def test_gpu_exact_mode_refuses_general_box():
    xy = inputs.uniform_points(500, 20.0, 20.0, seed=3)
    dg = DelaunayGPU(500, 16, 1.0, PeriodicBoundaries(20.0, 5.0, 0.0, 20.0))
    dg.setPrecision(PREC_EXACT)
    pts = torch.from_numpy(xy).cuda()
    tri = torch.zeros(500 * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(500, dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.GdtError) as e:
        dg.globalDelaunayTriangulation(pts, tri, num)
    assert e.value.status == _lib.GDT_ERR_INVALID
