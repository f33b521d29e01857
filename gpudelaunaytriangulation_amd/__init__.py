"""MI355X (gfx950) localized periodic Delaunay triangulation.

A drop-in for the hot path of DanielMSussman/gpuDelaunayTriangulation:
``DelaunayGPU.globalDelaunayTriangulation`` and
``DelaunayGPU.testAndRepairDelaunayTriangulation`` over a ``cellListGPU``,
implemented as hand-written HIP kernels behind the C ABI in
``include/gpudelaunay.h`` (``libgpudelaunay.so``, built in-tree by
``python -m gpudelaunaytriangulation_amd.build``).
"""
from . import inputs, rings
from ._lib import (GDT_OK, GDT_OVERFLOW, PREC_EXACT, PREC_FP32F, PREC_FP64, GdtError, LibraryNotBuilt,
                   load as load_library)

__all__ = [
    "inputs", "rings", "GDT_OK", "GDT_OVERFLOW", "PREC_EXACT", "PREC_FP32F", "PREC_FP64", "GdtError",
    "LibraryNotBuilt", "load_library", "DelaunayGPU", "CellListGPU", "PeriodicBoundaries",
    "Context", "BatchedDelaunayGPU",
]


def __getattr__(name):
    # torch-backed classes load lazily so the CPU-only tools import without torch
    if name in ("DelaunayGPU", "CellListGPU", "PeriodicBoundaries", "Context",
                "BatchedDelaunayGPU"):
        from . import delaunay
        return getattr(delaunay, name)
    raise AttributeError(name)
