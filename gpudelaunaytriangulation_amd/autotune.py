"""Autotuner for the triangulation's launch configuration.

Tunes what ``gdt_set_config`` exposes: threads per block of the one-ring
kernels (the reference fixes 128, DelaunayGPU.cu:16), the cell-size scale of
the internal grid (the reference takes the caller's cell size,
triangulation.cpp:172), the cell order (LDS tiles or row-major), the LDS
tile's edge (16x16 or 32x32 cells: the window every block stages) and the
lanes that share one point (one, two or four, for the global triangulation
and for repairs; the FP32-filtered kernel shares a point between at most
two; 64 for the global triangulation: the star kernel, one wavefront per
point).  Every candidate is timed on the caller's own points with HIP
events; the rings of every candidate must be identical (canonical hash) --
the output never depends on the configuration -- and the fastest is applied
to the context.  ``tune`` times global triangulations; ``tune_md`` times the
Brownian test-and-repair loop of a batch (hipGraph replays), where the
repairs' knobs matter.  Results are cached per workload and device for the
life of the process.
"""
from __future__ import annotations

import itertools

import torch

from . import _lib, rings
from .delaunay import DelaunayGPU

DEFAULT_BLOCKS = (256, 512)
DEFAULT_SCALES = (0.9, 1.0, 1.15)
KNOBS = ("block", "tiled", "cell_scale", "tile_shift", "lanes", "repair_lanes")

_cache: dict = {}


def _as_config(c) -> dict:
    """A candidate as a knob dict: a dict, or a (block, tiled, cell_scale
    [, tile_shift [, lanes [, repair_lanes]]]) tuple."""
    if isinstance(c, dict):
        return dict(c)
    return dict(zip(KNOBS, c))


def candidates(blocks=DEFAULT_BLOCKS, scales=DEFAULT_SCALES, tile_shifts=(4, 5),
               lanes=(1, 2, 4, 64), row_major=True):
    """Global-triangulation candidates: every block x scale x tile edge x
    lanes per point in tile order, plus the row-major order at the caller's
    cell size."""
    out = [dict(block=b, tiled=1, cell_scale=s, tile_shift=t, lanes=w)
           for b, s, t, w in itertools.product(blocks, scales, tile_shifts, lanes)]
    if row_major:
        out += [dict(block=b, tiled=0, cell_scale=1.0) for b in blocks]
    return out


def md_candidates(blocks=DEFAULT_BLOCKS, tile_shifts=(4, 5), repair_lanes=(1, 2, 4)):
    """Brownian-loop candidates: block x tile edge x lanes per repaired point."""
    return [dict(block=b, tiled=1, tile_shift=t, repair_lanes=r)
            for b, t, r in itertools.product(blocks, tile_shifts, repair_lanes)]


def _apply(ctx, cfg: dict, profile=None) -> None:
    ctx.set_config(profile=profile, **{k: cfg[k] for k in KNOBS if k in cfg})


def _restore(ctx, saved) -> None:
    ctx.set_config(block=saved.block, tiled=saved.tiled, cell_scale=saved.cell_scale,
                   profile=saved.profile, tile_shift=saved.tile_shift, lanes=saved.lanes,
                   repair_lanes=saved.repair_lanes)


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
    precision; apply and return the fastest (its knobs, ``ms``, ``table``,
    ``ring_hash``).  Raises if two configurations give different rings."""
    n = int(points.shape[0])
    x11, _, _, x22 = dg.Box.getBoxDims()
    dev = torch.cuda.get_device_name(points.device)
    key = ("tri", n, x11, x22, dg.cellsize, dg.precision, dev)
    if key in _cache:
        best = _cache[key]
        _apply(dg.ctx, best)
        return best
    configs = [_as_config(c) for c in (candidates() if configs is None else configs)]
    saved = dg.ctx.config()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device=points.device)
    num = torch.zeros(n, dtype=torch.int32, device=points.device)
    table, ref_hash = [], None
    try:
        for cfg in configs:
            _apply(dg.ctx, cfg, profile=0)
            ms = _time(dg, points, tri, num, reps)
            if check:
                nb = tri.cpu().numpy().reshape(n, dg.MaxSize)
                h = rings.ring_hash(nb, num.cpu().numpy())
                if ref_hash is None:
                    ref_hash = h
                elif h != ref_hash:
                    raise RuntimeError(f"configuration {cfg} changed the rings")
            table.append(dict(cfg, ms=ms))
    finally:
        _restore(dg.ctx, saved)
    best = min(table, key=lambda r: r["ms"])
    best = dict(best, table=table, ring_hash=ref_hash)
    _cache[key] = best
    _apply(dg.ctx, best)
    return best


def tune_md(bd, points, tri, num, configs=None, steps: int = 20, sigma: float = 0.01,
            seed: int = 1, check: bool = True) -> dict:
    """Tune the Brownian loop of the batch ``bd`` (a BatchedDelaunayGPU whose
    rings ``tri``/``num`` triangulate ``points``): every configuration runs
    ``steps`` hipGraph-replayed steps from the same start (a copy of the
    positions and rings), timed with HIP events; the final positions and
    rings must agree across configurations.  The caller's arrays are left as
    they were; the fastest configuration is applied and returned (its knobs,
    ``ms_per_step``, ``table``)."""
    dev = torch.cuda.get_device_name(points.device)
    key = ("md", tuple(bd.offsets), tuple(bd.boxes), bd.precision, steps, dev)
    if key in _cache:
        best = _cache[key]
        _apply(bd.ctx, best)
        return best
    configs = [_as_config(c) for c in (md_candidates() if configs is None else configs)]
    saved = bd.ctx.config()
    ms0 = bd.MaxSize
    table, ref = [], None
    try:
        for cfg in configs:
            _apply(bd.ctx, cfg, profile=0)
            bd.MaxSize = ms0
            X, T, C = points.clone(), tri.clone(), num.clone()
            bd.md_run(X, T, C, 2, sigma, seed=seed + 1, step0=0)       # warm-up: capture, widen
            bd.globalDelaunayTriangulation(X.copy_(points), T, C)      # back to the start
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            rep = bd.md_run(X, T, C, steps, sigma, seed=seed, step0=0, check_every=steps)
            e1.record()
            torch.cuda.synchronize()
            if check:
                nb = T.cpu().numpy().reshape(bd.N, bd.MaxSize)
                state = (X.cpu().numpy().tobytes(), rings.ring_hash(nb, C.cpu().numpy()))
                if ref is None:
                    ref = state
                elif state != ref:
                    raise RuntimeError(f"configuration {cfg} changed the Brownian run")
            table.append(dict(cfg, ms_per_step=e0.elapsed_time(e1) / steps,
                              recoveries=rep["recoveries"]))
    finally:
        _restore(bd.ctx, saved)
        bd.MaxSize = ms0
    best = min(table, key=lambda r: r["ms_per_step"])
    best = dict(best, table=table)
    _cache[key] = best
    _apply(bd.ctx, best)
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


__all__ = ["candidates", "md_candidates", "tune", "tune_md", "tuned_delaunay", "clear_cache"]
