"""GPU parity of the global triangulation (C ABI via the DelaunayGPU mirror)
against the CPU oracle and the golden fixtures."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpu_util import gpu_triangulate  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
PRECS = [PREC_FP64, PREC_FP32F]


@pytest.mark.parametrize("prec", PRECS)
def test_c1_matches_fixture(golden, prec):
    xy = inputs.uniform_points(4096, 64.0, 64.0, seed=1)
    nb, c, dg = gpu_triangulate(xy, 64.0, 64.0, prec)
    fx = np.load(os.path.join(HERE, "golden", "c1_n4096.npz"))
    assert np.array_equal(c, fx["counts"])
    assert np.array_equal(rings.canonical(nb, c), fx["canonical"])
    assert int(c.sum()) == 6 * 4096
    st = dg.stats()
    if prec == PREC_FP64:
        assert st["fallbacks"] == 0
    assert st["predicates"] > 0


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", ["c2", "c3b"])
def test_million_point_workloads_match_golden(golden, name, prec):
    w = golden["workloads"][name]
    if name == "c2":
        xy = inputs.uniform_points(1 << 20, 1024.0, 1024.0, seed=2)
    else:
        xy = inputs.lattice_points(1024, 1024, 1e-5, seed=3)
    assert inputs.sha256(xy) == w["points_sha256"]
    nb, c, _ = gpu_triangulate(xy, w["lx"], w["ly"], prec, maxn=16)
    assert int(c.sum()) == 6 * xy.shape[0]
    assert np.bincount(c).tolist() == w["degree_histogram"]
    assert rings.ring_hash(nb, c) == w["ring_sha256"]
    fx = np.load(os.path.join(HERE, "golden", f"{name}_samples.npz"))
    idx = fx["idx"]
    assert np.array_equal(rings.canonical(nb[idx], c[idx]), fx["canonical"])


@pytest.mark.parametrize("case", [("uniform65536", 65536, 256.0, 256.0),
                                  ("lattice256", None, *inputs.lattice_box(256, 256))])
def test_fp32_filtered_rows_identical_to_fp64(case):
    name, n, lx, ly = case
    xy = (inputs.uniform_points(n, lx, ly, seed=21) if n else
          inputs.lattice_points(256, 256, 1e-5, seed=22))
    n64, c64, dg = gpu_triangulate(xy, lx, ly, PREC_FP64)
    n32, c32, dg32 = gpu_triangulate(xy, lx, ly, PREC_FP32F)
    assert np.array_equal(c64, c32)
    assert np.array_equal(n64, n32)           # same start, same order, same padding
    st = dg32.stats()
    if name.startswith("lattice"):
        assert st["fallbacks"] > 0             # near-cocircular quads need FP64


@pytest.mark.parametrize("n,lx,ly,cell", [(1000, 40.0, 25.0, 1.0), (256, 16.0, 16.0, 1.0),
                                           (100, 10.0, 10.0, 1.0), (20000, 141.0, 141.0, 1.3),
                                           (5000, 70.0, 70.0, 0.7)])
@pytest.mark.parametrize("prec", PRECS)
def test_small_and_rectangular_match_oracle(n, lx, ly, cell, prec):
    xy = inputs.uniform_points(n, lx, ly, seed=n)
    on, oc = oracle.triangulate(xy, lx, ly, 1.0, 16)
    nb, c, _ = gpu_triangulate(xy, lx, ly, prec, cell=cell)
    assert rings.ring_hash(nb, c) == rings.ring_hash(on, oc)


def test_overflow_grows_max_size():
    """maxNeighbors = 4 (the reference minimum): the call grows MaxSize and
    resizes GPUTriangulation, then matches a roomy run."""
    xy = inputs.uniform_points(4096, 64.0, 64.0, seed=1)
    nb4, c4, dg = gpu_triangulate(xy, 64.0, 64.0, PREC_FP64, maxn=4)
    assert dg.MaxSize == c4.max() > 4 and nb4.shape[1] == dg.MaxSize
    nb, c, _ = gpu_triangulate(xy, 64.0, 64.0, PREC_FP64, maxn=16)
    assert rings.ring_hash(nb4, c4) == rings.ring_hash(nb, c)


def test_repeated_calls_are_bitwise_deterministic():
    xy = inputs.uniform_points(65536, 256.0, 256.0, seed=5)
    a = gpu_triangulate(xy, 256.0, 256.0, PREC_FP32F)
    b = gpu_triangulate(xy, 256.0, 256.0, PREC_FP32F, dg=a[2])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_row_major_and_tiled_cell_orders_agree():
    xy = inputs.uniform_points(20000, 141.0, 141.0, seed=8)
    nb_t, c_t, dg = gpu_triangulate(xy, 141.0, 141.0, PREC_FP64)
    dg.ctx.set_config(tiled=0, block=128)
    nb_r, c_r, _ = gpu_triangulate(xy, 141.0, 141.0, PREC_FP64, dg=dg)
    assert np.array_equal(nb_t, nb_r) and np.array_equal(c_t, c_r)


def test_morton_order_inside_tiles_agrees():
    xy = inputs.uniform_points(65536, 256.0, 256.0, seed=9)
    nb_t, c_t, dg = gpu_triangulate(xy, 256.0, 256.0, PREC_FP32F)
    dg.ctx.set_config(tiled=2)
    nb_m, c_m, _ = gpu_triangulate(xy, 256.0, 256.0, PREC_FP32F, dg=dg)
    assert dg.ctx.config().tiled == 2
    assert np.array_equal(nb_t, nb_m) and np.array_equal(c_t, c_m)


@pytest.mark.parametrize("prec", PRECS)
def test_point_order_does_not_change_the_rings(prec):
    """A permuted input gives the same triangulation after mapping indices back."""
    L = 256.0
    xy = inputs.uniform_points(65536, L, L, seed=10)
    perm = np.random.default_rng(11).permutation(xy.shape[0])
    nb, c, _ = gpu_triangulate(xy, L, L, prec)
    nbp, cp, _ = gpu_triangulate(xy[perm], L, L, prec)
    # permuted run: row j is original point perm[j] and its entries are
    # permuted indices; map entries through perm and rows through perm^-1
    back = np.where(nbp >= 0, perm[np.clip(nbp, 0, None)], -1)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    nb2, c2 = back[inv], cp[inv]
    assert np.array_equal(c2, c)
    assert np.array_equal(rings.canonical(nb2, c2), rings.canonical(nb, c))


@pytest.mark.parametrize("prec", PRECS)
def test_clustered_points_take_the_deferred_path(prec):
    """A dense cluster overflows the LDS window of its tiles; those points are
    deferred (wavefront kernel) and still match the oracle."""
    L = 141.0
    base = inputs.uniform_points(17000, L, L, seed=41)
    rng = np.random.default_rng(42)
    cluster = 60.0 + rng.uniform(0.0, 6.0, size=(3000, 2))
    xy = np.concatenate([base, cluster])
    on, oc = oracle.triangulate(xy, L, L, 1.0, 16)
    nb, c, dg = gpu_triangulate(xy, L, L, prec)
    assert rings.ring_hash(nb, c) == rings.ring_hash(on, oc)


@pytest.mark.parametrize("prec", PRECS)
def test_grid_smaller_than_the_halo(prec):
    """6x6 cells (search radius 2 < the 3-cell scan block): every point is
    deferred to the wavefront kernel, which triangulates them; the rings equal
    scipy's Qhull (the reference path itself keeps a virtual corner in such
    small boxes, and the oracle reproduces that)."""
    from qhull_periodic import periodic_rings
    xy = inputs.uniform_points(40, 6.0, 6.0, seed=43)
    qn, qc = periodic_rings(xy, 6.0, 6.0)
    nb, c, dg = gpu_triangulate(xy, 6.0, 6.0, prec)
    assert dg.stats()["deferred"] == 40
    assert rings.ring_hash(nb, c) == rings.ring_hash(qn, qc)
