"""ctypes binding of libgpudelaunay.so (the C ABI in include/gpudelaunay.h).

The library is built in-tree by ``gpudelaunaytriangulation_amd.build``.  There
is no fallback of any kind: if the shared library is missing or cannot be
loaded, every entry point raises ``LibraryNotBuilt``.  The library's CPU mode
(``gdt_host_*``, the reference's ``setGPUcompute(false)``) runs only when a
caller selects it; a device call never turns into a host one.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgpudelaunay.so")

GDT_OK = 0
GDT_OVERFLOW = 1
GDT_ERR_INVALID = -1
GDT_ERR_DEVICE = -2
GDT_ERR_NONGENERIC = -3
GDT_ERR_EULER = -4
GDT_ERR_SEARCH = -5
GDT_ERR_NOMEM = -6
GDT_ERR_SCRATCH = -7

PREC_FP64 = 64
PREC_FP32F = 32
PREC_EXACT = 128       # every decision exact (rectangular boxes)

# Every symbol include/gpudelaunay.h declares (checked by the CPU tests).
EXPORTS = (
    "gdt_create", "gdt_destroy", "gdt_set_config", "gdt_get_config", "gdt_get_stats",
    "gdt_status_string", "gdt_celllist_build", "gdt_celllist_info", "gdt_celllist_export",
    "gdt_triangulate", "gdt_triangles", "gdt_test_circumcircles", "gdt_repair",
    "gdt_test_and_repair", "gdt_voronoi_forward", "gdt_voronoi_adjoint",
    "gdt_voronoi_forward_batch", "gdt_voronoi_adjoint_batch", "gdt_triangulate_batch",
    "gdt_test_and_repair_batch", "gdt_md_run", "gdt_voronoi_vertices",
    "gdt_host_celllist", "gdt_host_triangulate", "gdt_host_test_and_repair",
    "gdt_host_voronoi_vertices", "gdt_triangulate_noalloc",
    "gdt_celllist_build_general", "gdt_triangulate_general", "gdt_test_and_repair_general",
    "gdt_md_run_general", "gdt_voronoi_vertices_general", "gdt_voronoi_forward_general",
    "gdt_voronoi_adjoint_general", "gdt_host_celllist_general", "gdt_host_triangulate_general",
    "gdt_host_test_and_repair_general", "gdt_host_voronoi_vertices_general",
)


class LibraryNotBuilt(RuntimeError):
    pass


class GdtError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {status_string(status)} ({status})")


class Config(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("tiled", ctypes.c_int), ("cell_scale", ctypes.c_double),
                ("profile", ctypes.c_int), ("tile_shift", ctypes.c_int), ("lanes", ctypes.c_int),
                ("repair_lanes", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("predicates", ctypes.c_ulonglong), ("fallbacks", ctypes.c_ulonglong),
        ("neighbours", ctypes.c_ulonglong), ("max_ring", ctypes.c_int), ("flagged", ctypes.c_int),
        ("xsize", ctypes.c_int), ("ysize", ctypes.c_int), ("bsx", ctypes.c_double),
        ("bsy", ctypes.c_double), ("ms_cells", ctypes.c_float), ("ms_ring", ctypes.c_float),
        ("deferred", ctypes.c_int), ("ms_defer", ctypes.c_float),
        ("profiled_calls", ctypes.c_ulonglong), ("ms_cells_total", ctypes.c_double),
        ("ms_ring_total", ctypes.c_double), ("ms_defer_total", ctypes.c_double),
        ("tile_shift", ctypes.c_int), ("lanes", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MdReport(ctypes.Structure):
    _fields_ = [("steps_done", ctypes.c_int), ("first_bad_step", ctypes.c_int),
                ("required", ctypes.c_int), ("max_flagged", ctypes.c_int),
                ("skipped_steps", ctypes.c_int), ("flagged_total", ctypes.c_longlong)]


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_ip = ctypes.POINTER(ctypes.c_int)


def _sig(L, name, args, res=ctypes.c_int):
    f = getattr(L, name)
    f.argtypes = args
    f.restype = res


def load():
    """Load the shared library (raises LibraryNotBuilt if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt(
            f"{LIB_PATH} not found: run `python -m gpudelaunaytriangulation_amd.build` "
            "(there is no CPU fallback)")
    # torch's HIP libraries need the unversioned libamdhip64.so from their own
    # directory: loaded after this library (which binds libamdhip64.so.7), they
    # would bring a second HIP runtime into the process, and this library's
    # runtime would find no device.  Loaded first, torch's runtime is the one
    # this library binds to.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise LibraryNotBuilt(f"cannot load {LIB_PATH}: {e}") from e
    _sig(L, "gdt_create", [ctypes.POINTER(_vp), _i])
    _sig(L, "gdt_destroy", [_vp], None)
    _sig(L, "gdt_set_config", [_vp, ctypes.POINTER(Config)])
    _sig(L, "gdt_get_config", [_vp, ctypes.POINTER(Config)])
    _sig(L, "gdt_get_stats", [_vp, ctypes.POINTER(Stats)])
    _sig(L, "gdt_status_string", [_i], ctypes.c_char_p)
    _sig(L, "gdt_celllist_build", [_vp, _vp, _i, _d, _d, _d, _vp])
    _sig(L, "gdt_celllist_info", [_vp, _ip, _ip, ctypes.POINTER(_d), ctypes.POINTER(_d), _ip, _vp])
    _sig(L, "gdt_celllist_export", [_vp, _vp, _vp, _i, _vp])
    _sig(L, "gdt_triangulate", [_vp, _vp, _i, _d, _d, _d, _i, _vp, _vp, _i, _ip, _vp])
    _sig(L, "gdt_triangles", [_vp, _vp, _vp, _i, _i, _vp, _i, _ip,
                              ctypes.POINTER(ctypes.c_longlong), _vp])
    _sig(L, "gdt_test_circumcircles", [_vp, _vp, _i, _vp, _i, _vp, _ip, _vp])
    _sig(L, "gdt_repair", [_vp, _vp, _i, _vp, _i, _vp, _vp, _i, _ip, _vp])
    _sig(L, "gdt_test_and_repair", [_vp, _vp, _i, _d, _d, _d, _i, _vp, _vp, _i, _ip, _ip, _vp])
    _sig(L, "gdt_voronoi_forward", [_vp, _vp, _i, _d, _d, _vp, _vp, _i, _d, _d, _i, _vp, _vp, _vp,
                                    _vp])
    _sig(L, "gdt_voronoi_adjoint", [_vp, _vp, _i, _d, _d, _vp, _vp, _i, _d, _d, _i, _vp, _vp, _vp,
                                    _vp, _vp])
    _sig(L, "gdt_voronoi_forward_batch", [_vp, _vp, _i, _ip, ctypes.POINTER(_d), _vp, _vp, _i, _d,
                                          _d, _i, _vp, _vp, _vp, _vp])
    _sig(L, "gdt_voronoi_adjoint_batch", [_vp, _vp, _i, _ip, ctypes.POINTER(_d), _vp, _vp, _i, _d,
                                          _d, _i, _vp, _vp, _vp, _vp, _vp])
    _sig(L, "gdt_triangulate_batch", [_vp, _vp, _i, _ip, ctypes.POINTER(_d), _d, _i, _vp, _vp, _i,
                                      _ip, _vp])
    _sig(L, "gdt_test_and_repair_batch", [_vp, _vp, _i, _ip, ctypes.POINTER(_d), _d, _i, _vp, _vp,
                                          _i, _ip, _ip, _vp])
    _sig(L, "gdt_voronoi_vertices", [_vp, _vp, _i, _d, _d, _vp, _vp, _i, _vp, _vp])
    _sig(L, "gdt_md_run", [_vp, _vp, _i, _ip, ctypes.POINTER(_d), _d, _i, _vp, _vp, _i, _d,
                           ctypes.c_ulonglong, _i, _i, _i, _i, ctypes.POINTER(MdReport), _vp])
    _sig(L, "gdt_triangulate_noalloc", [_vp, _i, _d, _d, _d, _i, _vp, _vp, _i, _ip, _vp])
    _sig(L, "gdt_host_celllist", [_vp, _i, _d, _d, _d, _vp, _vp, _i, _ip, _ip, _ip])
    _sig(L, "gdt_host_triangulate", [_vp, _i, _d, _d, _d, _i, _vp, _vp, _i, _i, _ip])
    _sig(L, "gdt_host_test_and_repair", [_vp, _i, _d, _d, _d, _i, _vp, _vp, _i, _i, _ip, _ip])
    _sig(L, "gdt_host_voronoi_vertices", [_vp, _i, _d, _d, _vp, _vp, _i, _vp, _i])
    _dp = ctypes.POINTER(_d)
    _sig(L, "gdt_celllist_build_general", [_vp, _vp, _i, _dp, _d, _vp])
    _sig(L, "gdt_triangulate_general", [_vp, _vp, _i, _ip, _dp, _d, _i, _vp, _vp, _i, _ip, _vp])
    _sig(L, "gdt_test_and_repair_general", [_vp, _vp, _i, _ip, _dp, _d, _i, _vp, _vp, _i, _ip,
                                            _ip, _vp])
    _sig(L, "gdt_md_run_general", [_vp, _vp, _i, _ip, _dp, _d, _i, _vp, _vp, _i, _d,
                                   ctypes.c_ulonglong, _i, _i, _i, _i, ctypes.POINTER(MdReport),
                                   _vp])
    _sig(L, "gdt_voronoi_vertices_general", [_vp, _vp, _i, _dp, _vp, _vp, _i, _vp, _vp])
    _sig(L, "gdt_voronoi_forward_general", [_vp, _vp, _i, _ip, _dp, _vp, _vp, _i, _d, _d, _i,
                                            _vp, _vp, _vp, _vp])
    _sig(L, "gdt_voronoi_adjoint_general", [_vp, _vp, _i, _ip, _dp, _vp, _vp, _i, _d, _d, _i,
                                            _vp, _vp, _vp, _vp, _vp])
    _sig(L, "gdt_host_celllist_general", [_vp, _i, _dp, _d, _vp, _vp, _i, _ip, _ip, _ip])
    _sig(L, "gdt_host_triangulate_general", [_vp, _i, _dp, _d, _i, _vp, _vp, _i, _i, _ip])
    _sig(L, "gdt_host_test_and_repair_general", [_vp, _i, _dp, _d, _i, _vp, _vp, _i, _i, _ip,
                                                 _ip])
    _sig(L, "gdt_host_voronoi_vertices_general", [_vp, _i, _dp, _vp, _vp, _i, _vp, _i])
    _lib = L
    return L


def status_string(st: int) -> str:
    try:
        return load().gdt_status_string(int(st)).decode()
    except LibraryNotBuilt:
        return "library not built"


def check(st: int, where: str, allow=(GDT_OK,)) -> int:
    if st not in allow:
        raise GdtError(st, where)
    return st
