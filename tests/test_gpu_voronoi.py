"""GPU Voronoi energy (forward) and its hand-written adjoint, through the C
ABI, against the oracle's FP64 energy and gradient (themselves pinned by
finite differences and an independent autograd restatement in
test_voronoi_oracle.py), plus torch.autograd.gradcheck of the
autograd.Function.

Workloads: 65,536 uniform points in a 256x256 box, and 64 sets of 4,096 points
in 64x64 boxes (batched).  Tolerances: FP64 relative 1e-11 on areas,
perimeters and energies, 1e-9 on gradients (scaled by the largest gradient
entry); FP32 path relative 5e-6 on areas and perimeters, 1e-7 on energies,
1e-4 on gradients (scaled the same way; measured at 65,536 points: 4.0e-7,
1.3e-9 and 7.1e-6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs  # noqa: E402
from gpudelaunaytriangulation_amd.delaunay import Context  # noqa: E402
from gpudelaunaytriangulation_amd.voronoi import (voronoi_adjoint, voronoi_energy,  # noqa: E402
                                                  voronoi_forward)

TOL = {PREC_FP64: dict(ap=1e-11, e=1e-11, g=1e-9), PREC_FP32F: dict(ap=5e-6, e=1e-7, g=1e-4)}


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def check_against_oracle(prec, area, perim, E, grad, o_E, o_area, o_perim, o_grad):
    t = TOL[prec]
    area, perim, grad = area.cpu().numpy(), perim.cpu().numpy(), grad.cpu().numpy()
    assert np.all(np.isfinite(area)) and np.all(np.isfinite(grad))
    assert np.abs(area - o_area).max() <= t["ap"] * np.abs(o_area).max()
    assert np.abs(perim - o_perim).max() <= t["ap"] * np.abs(o_perim).max()
    assert abs(E - o_E) <= t["e"] * abs(o_E)
    assert np.abs(grad - o_grad).max() <= t["g"] * np.abs(o_grad).max()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_single_set_matches_oracle(ctx, prec):
    L = 256.0
    xy = inputs.uniform_points(65536, L, L, seed=5)
    nb, c = oracle.triangulate(xy, L, L, 1.0, 16)
    o_E, o_area, o_perim, o_grad = oracle.voronoi_energy(xy, L, L, nb, c)
    X, NB, C = dev(xy, torch.float64), dev(nb, torch.int32), dev(c, torch.int32)
    area, perim, E = voronoi_forward(ctx, X, NB, C, (L, L), precision=prec)
    grad = voronoi_adjoint(ctx, X, NB, C, (L, L), area, perim, precision=prec)
    torch.cuda.synchronize()
    check_against_oracle(prec, area, perim, E.item(), grad, o_E, o_area, o_perim, o_grad)
    assert abs(area.sum().item() - L * L) <= 1e-9 * L * L


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32F])
def test_batch_matches_oracle_per_set(ctx, prec):
    nsets, n, L = 64, 4096, 64.0
    xs, nbs, cs, offs, ref = [], [], [], [0], []
    for s in range(nsets):
        xy = inputs.uniform_points(n, L, L, seed=1000 + s)
        nb, c = oracle.triangulate(xy, L, L, 1.0, 16)
        ref.append(oracle.voronoi_energy(xy, L, L, nb, c))
        xs.append(xy)
        nbs.append(np.where(nb >= 0, nb + offs[-1], nb))
        cs.append(c)
        offs.append(offs[-1] + n)
    X = dev(np.concatenate(xs), torch.float64)
    NB = dev(np.concatenate(nbs), torch.int32)
    C = dev(np.concatenate(cs), torch.int32)
    boxes = [(L, L)] * nsets
    gscale = torch.linspace(0.5, 2.0, nsets, dtype=torch.float64, device="cuda")
    area, perim, E = voronoi_forward(ctx, X, NB, C, boxes, offsets=offs, precision=prec)
    grad = voronoi_adjoint(ctx, X, NB, C, boxes, area, perim, gscale=gscale, offsets=offs,
                           precision=prec)
    torch.cuda.synchronize()
    E, g_s = E.cpu().numpy(), gscale.cpu().numpy()
    for s in range(nsets):
        sl = slice(offs[s], offs[s + 1])
        o_E, o_area, o_perim, o_grad = ref[s]
        check_against_oracle(prec, area[sl], perim[sl], E[s], grad[sl] / g_s[s], o_E, o_area,
                             o_perim, o_grad)


def test_autograd_function_gradcheck(ctx):
    L = 8.0
    xy = inputs.uniform_points(64, L, L, seed=7)
    nb, c = oracle.triangulate(xy, L, L, 1.0, 16)
    NB, C = dev(nb, torch.int32), dev(c, torch.int32)
    X = dev(xy, torch.float64).requires_grad_(True)

    def f(x):
        return voronoi_energy(x, NB, C, (L, L), context=ctx)

    assert torch.autograd.gradcheck(f, (X,), eps=1e-6, atol=1e-6, rtol=1e-5)
    # backward of a scaled loss uses the upstream gradient
    X.grad = None
    (3.0 * f(X)).sum().backward()
    g1 = X.grad.clone()
    X.grad = None
    f(X).sum().backward()
    # the scale enters before the per-slot sums: rounding differs at ~1e-15 of
    # the largest term, so compare against the largest entry
    assert (g1 - 3.0 * X.grad).abs().max() <= 1e-12 * g1.abs().max()


def test_rows_that_are_not_rings_give_nan(ctx):
    L = 16.0
    xy = inputs.uniform_points(256, L, L, seed=9)
    nb, c = oracle.triangulate(xy, L, L, 1.0, 16)
    nb[10, 0] = -5          # not a point index
    c[20] = 99              # longer than the stride
    X, NB, C = dev(xy, torch.float64), dev(nb, torch.int32), dev(c, torch.int32)
    area, perim, E = voronoi_forward(ctx, X, NB, C, (L, L))
    grad = voronoi_adjoint(ctx, X, NB, C, (L, L), area, perim)
    torch.cuda.synchronize()
    a, g = area.cpu().numpy(), grad.cpu().numpy()
    assert np.isnan(a[10]) and np.isnan(a[20]) and np.isnan(E.item())
    assert np.isnan(g[10]).all() and np.isnan(g[20]).all()
    ok = np.ones(256, bool)
    ok[[10, 20]] = False
    ok[nb[10, :c[10] if c[10] <= 16 else 16]] = False
    ok[nb[20]] = False
    assert np.isfinite(a[ok]).all()
