"""Helpers shared by the GPU tests (all calls go through the C ABI)."""
import numpy as np
import torch

from gpudelaunaytriangulation_amd import PREC_FP64
from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries


def gpu_triangulate(xy, lx, ly, prec=PREC_FP64, maxn=16, cell=1.0, dg=None):
    """(nbrs [n, MaxSize] int32 numpy, counts numpy, DelaunayGPU)."""
    n = xy.shape[0]
    if dg is None:
        dg = DelaunayGPU(n, maxn, cell, PeriodicBoundaries(lx, ly))
    dg.setPrecision(prec)
    pts = torch.from_numpy(np.ascontiguousarray(xy)).cuda()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    return tri.cpu().numpy().reshape(n, dg.MaxSize), num.cpu().numpy(), dg
