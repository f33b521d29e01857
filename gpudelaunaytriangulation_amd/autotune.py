"""Autotuner for the triangulation's launch configuration.

Tunes what ``gdt_set_config`` exposes: threads per block of the one-ring
kernels (the reference fixes 128, DelaunayGPU.cu:16), the cell-size scale of
the internal grid (the reference takes the caller's cell size,
triangulation.cpp:172) and the cell order (32x32-cell LDS tiles or
row-major).  Every candidate is timed on the caller's own points with HIP
events; the rings of every candidate must be identical (canonical hash) —
the output never depends on the configuration — and the fastest is applied
to the context.  Results are cached per (N, box, precision, device) for the
life of the process.
"""
from __future__ import annotations

import itertools

import torch

from . import _lib, rings
from .delaunay import DelaunayGPU

DEFAULT_BLOCKS = (128, 256, 320, 512)
DEFAULT_SCALES = (0.8, 0.9, 1.0, 1.15)

_cache: dict = {}


def candidates(blocks=DEFAULT_BLOCKS, scales=DEFAULT_SCALES, row_major=True):
    """(block, tiled, cell_scale) triples: every block x scale in tile order,
    plus the row-major order at the reference's cell size."""
    out = [(b, 1, s) for b, s in itertools.product(blocks, scales)]
    if row_major:
        out += [(b, 0, 1.0) for b in blocks]
    return out


def _time(dg, pts, tri, num, reps):
    dg.globalDelaunayTriangulation(pts, tri, num)          # warm-up (and growth)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dg.globalDelaunayTriangulation(pts, tri, num)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(dg: DelaunayGPU, points, configs=None, reps: int = 5, check: bool = True) -> dict:
    """Time every configuration on ``points`` with ``dg``'s box, cell size and
    precision; apply and return the fastest ({block, tiled, cell_scale, ms,
    table}).  Raises if two configurations give different rings."""
    n = int(points.shape[0])
    x11, _, _, x22 = dg.Box.getBoxDims()
    dev = torch.cuda.get_device_name(points.device)
    key = (n, x11, x22, dg.cellsize, dg.precision, dev)
    if key in _cache:
        best = _cache[key]
        dg.ctx.set_config(block=best["block"], tiled=best["tiled"], cell_scale=best["cell_scale"])
        return best
    configs = candidates() if configs is None else list(configs)
    saved = dg.ctx.config()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device=points.device)
    num = torch.zeros(n, dtype=torch.int32, device=points.device)
    table, ref_hash = [], None
    try:
        for block, tiled, scale in configs:
            dg.ctx.set_config(block=block, tiled=tiled, cell_scale=scale, profile=0)
            ms = _time(dg, points, tri, num, reps)
            if check:
                nb = tri.cpu().numpy().reshape(n, dg.MaxSize)
                h = rings.ring_hash(nb, num.cpu().numpy())
                if ref_hash is None:
                    ref_hash = h
                elif h != ref_hash:
                    raise RuntimeError(f"configuration {(block, tiled, scale)} changed the rings")
            table.append(dict(block=block, tiled=tiled, cell_scale=scale, ms=ms))
    finally:
        dg.ctx.set_config(block=saved.block, tiled=saved.tiled, cell_scale=saved.cell_scale,
                          profile=saved.profile)
    best = min(table, key=lambda r: r["ms"])
    best = dict(best, table=table, ring_hash=ref_hash)
    _cache[key] = best
    dg.ctx.set_config(block=best["block"], tiled=best["tiled"], cell_scale=best["cell_scale"])
    return best


def tuned_delaunay(points, box, maximumNeighborsGuess: int = 16, cellSize: float = 1.0,
                   precision: int = _lib.PREC_FP64, **kw) -> tuple[DelaunayGPU, dict]:
    """A DelaunayGPU for ``points`` with its configuration tuned."""
    from .delaunay import PeriodicBoundaries
    dg = DelaunayGPU(int(points.shape[0]), maximumNeighborsGuess, cellSize,
                     PeriodicBoundaries(*box))
    dg.setPrecision(precision)
    return dg, tune(dg, points, **kw)


def clear_cache() -> None:
    _cache.clear()


__all__ = ["candidates", "tune", "tuned_delaunay", "clear_cache"]
