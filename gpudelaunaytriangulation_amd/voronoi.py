"""Voronoi energy of a periodic point set and its gradient, on the gfx950
kernels behind ``gdt_voronoi_forward`` / ``gdt_voronoi_adjoint``.

    E = sum_i (A_i - a0)^2 + (P_i - p0)^2

with A_i, P_i the area and perimeter of the Voronoi cell of point i, whose
vertices are the circumcentres of the Delaunay triangles around i (the
vertices the reference keeps per point in ``GPUVoroCur``,
src/model/DelaunayGPU.cu:921-922).  The reference has no energy; this is the
differentiable layer a vertex-model caller builds on it.

``VoronoiEnergy`` is a ``torch.autograd.Function``: its backward is the
library's hand-written adjoint (no automatic differentiation of the
kernels).  The ring table is treated as constant: the energy is a smooth
function of the positions while the triangulation does not change.

Batched form: several sets concatenated along dim 0, ``offsets`` (host ints,
nsets+1) and ``boxes`` [(lx, ly)] per set; ring entries index the
concatenated array and the energy is one value per set.  A box may also be a
box matrix (x11, x12, x21, x22) with the lattice vectors as columns (a
general periodic box; minimum images by rounded unit-cell coordinates).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .delaunay import Context, _box_tuple, _dev_ptr, _stream_ptr

_D = ctypes.c_double


def _check_rings(xy, nbrs, counts):
    n = int(xy.shape[0])
    if xy.dim() != 2 or xy.shape[1] != 2:
        raise ValueError("points must be [N, 2]")
    if counts.numel() != n:
        raise ValueError("counts must have N entries")
    if nbrs.numel() % n != 0:
        raise ValueError("ring table must be [N, stride]")
    return n, nbrs.numel() // n


def _sets(n, offsets, boxes):
    """(nsets, offsets, boxes, general): (lx, ly) pairs when every box is a
    rectangle, else box matrices (4 doubles per set) for the general calls."""
    offs = [int(o) for o in offsets]
    if len(offs) != len(boxes) + 1 or offs[0] != 0 or offs[-1] != n:
        raise ValueError("offsets must run from 0 to N with one box per set")
    off_arr = (ctypes.c_int * len(offs))(*offs)
    tb = [_box_tuple(b) for b in boxes]
    general = not all(len(b) == 2 for b in tb)
    if general:
        flat = [v for b in tb for v in ((b[0], 0.0, 0.0, b[1]) if len(b) == 2 else b)]
    else:
        flat = [v for b in tb for v in b]
    box_arr = (_D * len(flat))(*flat)
    return len(boxes), off_arr, box_arr, general


def voronoi_forward(ctx: Context, xy, nbrs, counts, boxes, offsets=None, a0=1.0, p0=3.8,
                    precision=_lib.PREC_FP64):
    """Areas [N], perimeters [N] and energies [nsets] (float64 tensors)."""
    n, stride = _check_rings(xy, nbrs, counts)
    area = torch.empty(n, dtype=torch.float64, device=xy.device)
    perim = torch.empty_like(area)
    L = _lib.load()
    if offsets is None and len(_box_tuple(boxes)) == 4:
        offsets, boxes = [0, n], [boxes]
    if offsets is None:
        lx, ly = _box_tuple(boxes)
        energy = torch.empty(1, dtype=torch.float64, device=xy.device)
        st = L.gdt_voronoi_forward(ctx.handle, _dev_ptr(xy, torch.float64, "points"), n, float(lx),
                                   float(ly), _dev_ptr(nbrs, torch.int32, "rings"),
                                   _dev_ptr(counts, torch.int32, "counts"), stride, float(a0),
                                   float(p0), int(precision), area.data_ptr(), perim.data_ptr(),
                                   energy.data_ptr(), _stream_ptr())
    else:
        nsets, off_arr, box_arr, general = _sets(n, offsets, boxes)
        energy = torch.empty(nsets, dtype=torch.float64, device=xy.device)
        fn = L.gdt_voronoi_forward_general if general else L.gdt_voronoi_forward_batch
        st = fn(ctx.handle, _dev_ptr(xy, torch.float64, "points"), nsets,
                                         off_arr, box_arr, _dev_ptr(nbrs, torch.int32, "rings"),
                                         _dev_ptr(counts, torch.int32, "counts"), stride, float(a0),
                                         float(p0), int(precision), area.data_ptr(),
                                         perim.data_ptr(), energy.data_ptr(), _stream_ptr())
    _lib.check(st, "gdt_voronoi_forward")
    return area, perim, energy


def voronoi_adjoint(ctx: Context, xy, nbrs, counts, boxes, area, perim, gscale=None, offsets=None,
                    a0=1.0, p0=3.8, precision=_lib.PREC_FP64):
    """gscale[s] * dE_s/dx for every point (float64 [N, 2])."""
    n, stride = _check_rings(xy, nbrs, counts)
    grad = torch.empty(n, 2, dtype=torch.float64, device=xy.device)
    gs = 0
    if gscale is not None:
        gscale = gscale.detach().to(torch.float64).contiguous()
        gs = _dev_ptr(gscale, torch.float64, "gscale")
    L = _lib.load()
    if offsets is None and len(_box_tuple(boxes)) == 4:
        offsets, boxes = [0, n], [boxes]
    if offsets is None:
        lx, ly = _box_tuple(boxes)
        st = L.gdt_voronoi_adjoint(ctx.handle, _dev_ptr(xy, torch.float64, "points"), n, float(lx),
                                   float(ly), _dev_ptr(nbrs, torch.int32, "rings"),
                                   _dev_ptr(counts, torch.int32, "counts"), stride, float(a0),
                                   float(p0), int(precision), area.data_ptr(), perim.data_ptr(),
                                   gs or None, grad.data_ptr(), _stream_ptr())
    else:
        nsets, off_arr, box_arr, general = _sets(n, offsets, boxes)
        fn = L.gdt_voronoi_adjoint_general if general else L.gdt_voronoi_adjoint_batch
        st = fn(ctx.handle, _dev_ptr(xy, torch.float64, "points"), nsets,
                                         off_arr, box_arr, _dev_ptr(nbrs, torch.int32, "rings"),
                                         _dev_ptr(counts, torch.int32, "counts"), stride, float(a0),
                                         float(p0), int(precision), area.data_ptr(),
                                         perim.data_ptr(), gs or None, grad.data_ptr(),
                                         _stream_ptr())
    _lib.check(st, "gdt_voronoi_adjoint")
    return grad


class VoronoiEnergy(torch.autograd.Function):
    """E(xy) for a fixed ring table; backward = the library's adjoint.

    ``VoronoiEnergy.apply(xy, nbrs, counts, boxes, offsets, a0, p0, precision, ctx)``
    returns a float64 tensor of shape [1] (single set, ``offsets=None``) or
    [nsets]."""

    @staticmethod
    def forward(fctx, xy, nbrs, counts, boxes, offsets, a0, p0, precision, context):
        xy_c = xy.detach().contiguous()
        area, perim, energy = voronoi_forward(context, xy_c, nbrs, counts, boxes, offsets, a0, p0,
                                              precision)
        fctx.save_for_backward(xy_c, nbrs, counts, area, perim)
        fctx.args = (boxes, offsets, a0, p0, precision, context)
        return energy

    @staticmethod
    def backward(fctx, grad_energy):
        xy_c, nbrs, counts, area, perim = fctx.saved_tensors
        boxes, offsets, a0, p0, precision, context = fctx.args
        grad = voronoi_adjoint(context, xy_c, nbrs, counts, boxes, area, perim,
                               gscale=grad_energy, offsets=offsets, a0=a0, p0=p0,
                               precision=precision)
        return grad.to(xy_c.dtype), None, None, None, None, None, None, None, None


def voronoi_energy(xy, nbrs, counts, boxes, offsets=None, a0=1.0, p0=3.8,
                   precision=_lib.PREC_FP64, context: Context | None = None):
    """Differentiable energy (see VoronoiEnergy)."""
    if context is None:
        context = Context(xy.device.index or 0)
    return VoronoiEnergy.apply(xy, nbrs, counts, boxes, offsets, a0, p0, precision, context)
