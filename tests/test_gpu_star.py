"""The star kernel (lanes = 64: one wavefront per point, every candidate's
successor searched at once, the ring read off the lanes) gives the rings of
the sequential tile kernel and of the oracle: uniform points at the small sizes
it serves and at 2^18, a rectangular box, both tile edges, a cluster whose
blocks hold more than 64 entries and whose windows overflow (sequential steps
and deferral), a jittered lattice (near-cocircular quads: uncertified
successors), exact square lattices (exact ties, decided by the symbolic
perturbation after deferral), denser cells and a clump (the lanes' 5x5
blocks and lane 0's sequential steps) and batches of a few small sets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from exact_delaunay import SCALE, square_lattice  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries  # noqa: E402

STAR = 64


def triangulate(xy, lx, ly, prec=PREC_FP64, lanes=STAR, ts=0, cell=1.0):
    n = xy.shape[0]
    dg = DelaunayGPU(n, 16, cell, PeriodicBoundaries(lx, ly))
    dg.setPrecision(prec)
    dg.ctx.set_config(lanes=lanes, tile_shift=ts)
    pts = torch.from_numpy(np.ascontiguousarray(xy)).cuda()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    nb = tri.cpu().numpy().reshape(n, dg.MaxSize)
    return nb, num.cpu().numpy(), dg


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
@pytest.mark.parametrize("n,lx,ly,seed,ts", [(4096, 64.0, 64.0, 1, 0), (16384, 128.0, 128.0, 5, 0),
                                             (65536, 256.0, 256.0, 61, 0),
                                             (20000, 200.0, 100.0, 62, 0),
                                             (16384, 128.0, 128.0, 6, 5),
                                             (1 << 18, 512.0, 512.0, 7, 0)])
def test_star_equals_oracle(n, lx, ly, seed, ts, prec):
    print(f"[star] n {n} box {lx}x{ly} ts {ts} prec {prec}", flush=True)
    xy = inputs.uniform_points(n, lx, ly, seed=seed)
    nb, cnt, dg = triangulate(xy, lx, ly, prec, ts=ts)
    assert dg.stats()["lanes"] == STAR
    on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(on, oc)
    assert dg.stats()["neighbours"] == int(oc.sum())


def test_star_cluster_runs_sequential_steps_and_defers():
    """3,000 points in a 6x6 square: cell blocks of hundreds of entries (the
    lanes cannot hold them) and windows beyond the LDS budget (deferred)."""
    L = 141.0
    base = inputs.uniform_points(17000, L, L, seed=41)
    cluster = 60.0 + np.random.default_rng(42).uniform(0.0, 6.0, size=(3000, 2))
    xy = np.concatenate([base, cluster])
    on, oc = oracle.triangulate(xy, L, L, 1.0, 16)
    for ts in (4, 5):
        print(f"[star] cluster ts {ts}", flush=True)
        nb, cnt, dg = triangulate(xy, L, L, PREC_FP64, ts=ts)
        if ts == 4:
            assert dg.stats()["deferred"] > 0
        assert rings.ring_hash(nb, cnt) == rings.ring_hash(on, oc)


@pytest.mark.parametrize("case", ["coarse", "bump"])
def test_star_large_blocks_equal_oracle(case):
    """coarse: 1.44 points per cell (cell 1.2), the lanes' 7x7 blocks mostly
    exceed 64 entries (their 5x5 blocks); bump: a clump of 300 points in a 2x2
    square, where even the 5x5 blocks do (the sequential steps in lane 0);
    16x16-cell tiles, whose windows still fit."""
    n, L = 16384, 128.0
    xy = inputs.uniform_points(n, L, L, seed=9)
    cell = 1.2 if case == "coarse" else 1.0
    if case == "bump":
        xy = np.concatenate([xy, 40.0 + 2.0 * inputs.uniform_points(300, 1.0, 1.0, seed=10)])
    print(f"[star] large blocks: {case}", flush=True)
    nb, cnt, dg = triangulate(xy, L, L, PREC_FP64, ts=4, cell=cell)
    assert dg.stats()["lanes"] == STAR
    on, oc = oracle.triangulate(xy, L, L, 1.0, 16)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(on, oc)


def test_star_jittered_lattice_equals_oracle():
    lx, ly = inputs.lattice_box(128, 128)
    xy = inputs.lattice_points(128, 128)
    print("[star] jittered lattice 128x128", flush=True)
    nb, cnt, _ = triangulate(xy, lx, ly, PREC_FP64)
    on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(on, oc)


@pytest.mark.parametrize("nx,ny,sp", [(64, 64, 1.0), (96, 96, 0.25)])
def test_star_exact_lattice_equals_sequential(nx, ny, sp):
    """Exact ties: the rings equal the sequential kernel's (which the exact
    tie tests pin to the perturbed Delaunay triangulation)."""
    P = square_lattice(nx, ny, sp)
    xy, lx, ly = P / SCALE, nx * sp, ny * sp
    print(f"[star] exact lattice {nx}x{ny} spacing {sp}", flush=True)
    nb, cnt, _ = triangulate(xy, lx, ly, PREC_FP64)
    sb, sc, _ = triangulate(xy, lx, ly, PREC_FP64, lanes=1)
    assert int(cnt.sum()) == 6 * len(cnt)
    assert rings.ring_hash(nb, cnt) == rings.ring_hash(sb, sc)


def test_star_small_batches_equal_oracle_per_set():
    """A few small sets (the automatic choice then picks the star kernel; the
    blocks take their set from the tile they serve), rectangles among them,
    each set's rings equal the oracle's for that set alone."""
    from gpudelaunaytriangulation_amd.delaunay import BatchedDelaunayGPU
    specs = [(500, 25.0, 20.0, 81), (1200, 30.0, 40.0, 82), (3000, 55.0, 55.0, 83), (64, 8.0, 8.0, 84)]
    pts = [inputs.uniform_points(n, lx, ly, seed=s) for n, lx, ly, s in specs]
    for lanes in (0, STAR):
        bd = BatchedDelaunayGPU([n for n, *_ in specs], [(lx, ly) for _, lx, ly, _ in specs], 16, 1.0,
                                precision=PREC_FP64)
        bd.ctx.set_config(lanes=lanes)
        X = torch.from_numpy(np.concatenate(pts)).cuda()
        tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
        num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
        print(f"[star] small batch lanes {lanes}", flush=True)
        bd.globalDelaunayTriangulation(X, tri, num)
        torch.cuda.synchronize()
        assert bd.stats()["lanes"] == STAR
        nb = tri.cpu().numpy().reshape(bd.N, bd.MaxSize)
        cnt = num.cpu().numpy()
        for s, (xy, (n, lx, ly, _)) in enumerate(zip(pts, specs)):
            a, b = bd.offsets[s], bd.offsets[s + 1]
            local = nb[a:b].copy()
            local[local >= 0] -= a
            on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
            assert rings.ring_hash(local, cnt[a:b]) == rings.ring_hash(on, oc), s
