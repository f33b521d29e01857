"""Python mirror of the reference's host API over the gfx950 C ABI.

Names, arguments and error behaviour follow the reference classes
``DelaunayGPU`` (src/model/DelaunayGPU.h:18-130), ``cellListGPU``
(src/utilities/cellListGPU.h:17-121), ``periodicBoundaries``
(src/utilities/periodicBoundaries.h:25-73) and ``multiProfiler``
(src/utilities/multiProfiler.h:9-75).  Arrays are torch tensors on a HIP
device, standing in for the reference's ``GPUArray``:

* points            float64 [N, 2]        (GPUArray<double2>)
* GPUTriangulation  int32   [N*MaxSize]   (GPUArray<int>, read as [N][MaxSize])
* cellNeighborNum   int32   [N]

As in the reference, the triangulation call may grow ``MaxSize`` and resize
the caller's ``GPUTriangulation`` in place (DelaunayGPU.cpp:130, 154).
Unlike the reference, ring overflow is always detected, whatever
``setSafetyMode`` says (the reference silently truncates when it is off,
DelaunayGPU.cu:874-879).

CPU mode (``setGPUcompute(False)``, with ``setOMPthreads``) runs the library's
own per-point routines on the host (``gdt_host_*``), on CPU tensors, as the
reference's OpenMP branches do; it never stands in for a device call (a
device call with CPU tensors raises).
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def _stream_ptr(stream=None) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _dev_ptr(t, dtype, name: str) -> int:
    if torch is None or not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on a HIP device")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return int(t.data_ptr())


class PeriodicBoundaries:
    """Periodic box (periodicBoundaries, periodicBoundaries.h:25-73): a
    rectangle (``PeriodicBoundaries(x, y)``, setSquare, 75-84) or a general
    box (``PeriodicBoundaries(a, b, c, d)``, setGeneral, 86-101) whose matrix
    ``[[a, b], [c, d]]`` has the lattice vectors as columns."""

    def __init__(self, x: float = 1.0, y: float | None = None, c: float | None = None,
                 d: float | None = None):
        if c is None and d is None:
            self.setSquare(x, x if y is None else y)
        else:
            self.setGeneral(x, 0.0 if y is None else y, 0.0 if c is None else c,
                            1.0 if d is None else d)

    def setSquare(self, x: float, y: float) -> None:
        if not (x > 0 and y > 0):
            raise ValueError("box sides must be positive")
        self.x11, self.x12, self.x21, self.x22 = float(x), 0.0, 0.0, float(y)
        self._inverse()
        self.isSquare = True

    def setGeneral(self, a: float, b: float, c: float, d: float) -> None:
        m = [float(a), float(b), float(c), float(d)]
        det = m[0] * m[3] - m[1] * m[2]
        if not (all(math.isfinite(v) for v in m) and math.isfinite(det) and det != 0.0):
            raise ValueError("the box matrix must be finite and invertible")
        self.x11, self.x12, self.x21, self.x22 = m
        self._inverse()
        self.isSquare = False

    def _inverse(self) -> None:
        det = self.x11 * self.x22 - self.x12 * self.x21
        self.xi11, self.xi12 = self.x22 / det, -self.x12 / det
        self.xi21, self.xi22 = -self.x21 / det, self.x11 / det

    def getBoxDims(self):
        return self.x11, self.x12, self.x21, self.x22

    def getBoxInvDims(self):
        return self.xi11, self.xi12, self.xi21, self.xi22

    def isBoxSquare(self) -> bool:
        return self.isSquare

    def is_rectangle(self) -> bool:
        """Whether the library runs this box on its rectangular path
        (off-diagonal entries zero, positive sides)."""
        return self.x12 == 0.0 and self.x21 == 0.0 and self.x11 > 0 and self.x22 > 0

    def box_matrix(self) -> np.ndarray:
        """(x11, x12, x21, x22) as the general-box C calls take it."""
        return np.array([self.x11, self.x12, self.x21, self.x22], dtype=np.float64)

    def Trans(self, p):
        """Unit-cell coordinates to the box (periodicBoundaries.h:103-115)."""
        p = np.asarray(p, dtype=np.float64)
        return np.stack([self.x11 * p[..., 0] + self.x12 * p[..., 1],
                         self.x21 * p[..., 0] + self.x22 * p[..., 1]], -1)

    def invTrans(self, p):
        """Box to unit-cell coordinates (periodicBoundaries.h:117-129)."""
        p = np.asarray(p, dtype=np.float64)
        return np.stack([self.xi11 * p[..., 0] + self.xi12 * p[..., 1],
                         self.xi21 * p[..., 0] + self.xi22 * p[..., 1]], -1)

    def minDist(self, p1, p2):
        """Nearest-image displacement p1 - p2 (periodicBoundaries.h:154-179):
        rectangles per axis, general boxes on unit-cell coordinates wrapped
        into [-1/2, 1/2]."""
        if self.isSquare:
            d = np.asarray(p1, dtype=np.float64) - np.asarray(p2, dtype=np.float64)
            L = np.array([self.x11, self.x22])
            return d - L * np.round(d / L)
        u = self.invTrans(p1) - self.invTrans(p2)
        u = u - np.round(u)
        return self.Trans(u)

    def putInBoxReal(self, p):
        """Wrap into the unit cell (periodicBoundaries.h:131-152)."""
        p = np.asarray(p, dtype=np.float64)
        if self.isSquare:
            L = np.array([self.x11, self.x22])
            q = p - np.floor(p / L) * L
            return np.where(q >= L, q - L, q)
        u = self.invTrans(p)
        u = u - np.floor(u)
        u = np.where(u >= 1.0, u - 1.0, u)
        return self.Trans(u)

    def move(self, p, disp):
        """p + disp, put back in the box (periodicBoundaries.h:181-189)."""
        return self.putInBoxReal(np.asarray(p, dtype=np.float64) + np.asarray(disp, dtype=np.float64))


PeriodicBoxPtr = PeriodicBoundaries


def _box_arg(box) -> ctypes.Array:
    """A host double array of box matrices (4 doubles per box) for the C calls."""
    flat = np.asarray(box, dtype=np.float64).reshape(-1)
    return (ctypes.c_double * len(flat))(*flat.tolist())


class MultiProfiler:
    """Named wall-clock timers (multiProfiler, multiProfiler.h:9-75): ``start``
    and ``end`` bracket a region, ``timing`` is its mean duration in seconds
    over the completed calls, ``print`` lists every region."""

    def __init__(self):
        self.numberOfProfiles = 0
        self.processMap: dict[str, int] = {}
        self.startTimes: list[float] = []
        self.endTimes: list[float] = []
        self.functionCalls: list[int] = []
        self.timeTaken: list[float] = []

    def addName(self, name: str) -> None:
        if name not in self.processMap:
            self.processMap[name] = self.numberOfProfiles
            self.startTimes.append(0.0)
            self.endTimes.append(0.0)
            self.functionCalls.append(0)
            self.timeTaken.append(0.0)
            self.numberOfProfiles += 1

    def start(self, name: str) -> None:
        self.addName(name)
        self.startTimes[self.processMap[name]] = time.perf_counter()

    def end(self, name: str) -> None:
        i = self.processMap[name]
        self.endTimes[i] = time.perf_counter()
        self.timeTaken[i] += self.endTimes[i] - self.startTimes[i]
        self.functionCalls[i] += 1

    def timing(self, name: str) -> float:
        i = self.processMap.get(name)
        if i is None or self.functionCalls[i] == 0:
            return 0.0
        return self.timeTaken[i] / self.functionCalls[i]

    def print(self) -> None:
        for name in sorted(self.processMap):
            print(f"process {name} averaged\t {self.timing(name)} \t over "
                  f"{self.functionCalls[self.processMap[name]]} function calls")


multiProfiler = MultiProfiler


def _host_ptr(t, dtype, name: str) -> int:
    """Address of a contiguous CPU tensor (CPU mode's stand-in for a host handle)."""
    if torch is None or not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.device.type != "cpu":
        raise ValueError(f"{name} must be a CPU tensor in CPU mode (setGPUcompute(False))")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return int(t.data_ptr())


class Context:
    """Owns one gdt_context (device scratch buffers) on a HIP device."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.gdt_create(ctypes.byref(h), int(device)), "gdt_create")
        self._h = h
        self.device = int(device)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().gdt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self) -> _lib.Config:
        c = _lib.Config()
        _lib.check(_lib.load().gdt_get_config(self._h, ctypes.byref(c)), "gdt_get_config")
        return c

    def set_config(self, block: int | None = None, tiled: int | None = None,
                   cell_scale: float | None = None, profile: int | None = None,
                   tile_shift: int | None = None, lanes: int | None = None,
                   repair_lanes: int | None = None) -> None:
        c = self.config()
        if tile_shift is not None:
            c.tile_shift = int(tile_shift)
        if lanes is not None:
            c.lanes = int(lanes)
        if repair_lanes is not None:
            c.repair_lanes = int(repair_lanes)
        if profile is not None:
            c.profile = int(profile)
        if block is not None:
            c.block = int(block)
        if tiled is not None:
            c.tiled = int(tiled)
        if cell_scale is not None:
            c.cell_scale = float(cell_scale)
        _lib.check(_lib.load().gdt_set_config(self._h, ctypes.byref(c)), "gdt_set_config")

    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(_lib.load().gdt_get_stats(self._h, ctypes.byref(s)), "gdt_get_stats")
        return s.as_dict()


class CellListGPU:
    """Cell list over a periodic box (cellListGPU, cellListGPU.h:17-121).

    The list is a counting sort (cell-ordered arrays); ``cell_sizes`` and the
    padded ``idxs`` of the reference layout are produced on request."""

    def __init__(self, box: PeriodicBoundaries | None = None, context=None, device: int = 0):
        self.Box = box if box is not None else PeriodicBoundaries(1.0)
        # a Context, or a callable giving one; made on first device use
        self._ctx_src = context
        self._ctx = context if isinstance(context, Context) else None
        self.device = int(device)
        self.Np = 0
        self.a = 1.0
        self.xsize = self.ysize = 0
        self.boxsize = 0.0
        self.Nmax = 0
        self._cell_sizes = None
        self._idxs = None

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = self._ctx_src() if callable(self._ctx_src) else Context(self.device)
        return self._ctx

    def setNp(self, nn: int) -> None:
        self.Np = int(nn)

    def setBox(self, bx: PeriodicBoundaries) -> None:
        self.Box = bx

    def setGridSize(self, a: float) -> None:
        """cellListGPU::setGridSize (cellListGPU.cpp:97-117).  A general box's
        grid (cells of its reduced lattice basis) is known after compute()."""
        x11, _, _, x22 = self.Box.getBoxDims()
        self.a = float(a)
        if not self.Box.is_rectangle():
            self.xsize = self.ysize = 0
            self.boxsize = 0.0
            return
        self.xsize = int(math.floor(x11 / a))
        self.xsize += self.xsize % 2
        self.ysize = int(math.floor(x22 / a))
        self.ysize += self.ysize % 2
        self.boxsize = x11 / self.xsize

    def getNmax(self) -> int:
        return self.Nmax

    def getXsize(self) -> int:
        return self.xsize

    def getYsize(self) -> int:
        return self.ysize

    def getBoxsize(self) -> float:
        return self.boxsize

    def positionToCellIndex(self, x: float, y: float) -> int:
        """cellListGPU.cpp:183-189 (rectangular boxes, as the reference's)."""
        if not self.Box.is_rectangle():
            raise ValueError("positionToCellIndex is defined for rectangular boxes; a general "
                             "box's cells follow its reduced lattice basis")
        bsy = self.Box.getBoxDims()[3] / self.ysize
        bx = max(0, min(self.xsize - 1, int(math.floor(x / self.boxsize))))
        by = max(0, min(self.ysize - 1, int(math.floor(y / bsy))))
        return by * self.xsize + bx

    def getCellNeighbors(self, cellIndex: int, width: int) -> list[int]:
        """All cells within `width` of a cell (cellListGPU.cpp:196-212, row index
        taken as cellIndex // xsize)."""
        w = min(width, self.xsize // 2)
        cx, cy = cellIndex % self.xsize, cellIndex // self.xsize
        return [((cy + i) % self.ysize) * self.xsize + (cx + j) % self.xsize
                for i in range(-w, w + 1) for j in range(-w, w + 1)]

    def getCellShellNeighbors(self, cellIndex: int, width: int) -> list[int]:
        """Cells on the square shell at distance `width` (cellListGPU.cpp:221-237)."""
        w = min(width, self.xsize)
        cx, cy = cellIndex % self.xsize, cellIndex // self.xsize
        return [((cy + i) % self.ysize) * self.xsize + (cx + j) % self.xsize
                for i in range(-w, w + 1) for j in range(-w, w + 1)
                if i in (-w, w) or j in (-w, w)]

    def computeGPU(self, points, stream=None) -> None:
        """Bin the points (cellListGPU::computeGPU(points), cellListGPU.cpp:404-460)."""
        L = _lib.load()
        n = int(points.shape[0])
        x11, _, _, x22 = self.Box.getBoxDims()
        if self.xsize == 0:
            self.setGridSize(self.a)
        ptr = _dev_ptr(points, torch.float64, "points")
        if self.Box.is_rectangle():
            _lib.check(L.gdt_celllist_build(self.ctx.handle, ptr, n, x11, x22, self.a,
                                            _stream_ptr(stream)), "gdt_celllist_build")
        else:
            _lib.check(L.gdt_celllist_build_general(self.ctx.handle, ptr, n,
                                                    _box_arg(self.Box.box_matrix()), self.a,
                                                    _stream_ptr(stream)),
                       "gdt_celllist_build_general")
        self.Np = n
        xs, ys, mo = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        bx, by = ctypes.c_double(), ctypes.c_double()
        _lib.check(L.gdt_celllist_info(self.ctx.handle, ctypes.byref(xs), ctypes.byref(ys),
                                       ctypes.byref(bx), ctypes.byref(by), ctypes.byref(mo),
                                       _stream_ptr(stream)), "gdt_celllist_info")
        self.xsize, self.ysize, self.boxsize = xs.value, ys.value, bx.value
        self.Nmax = max(self.Nmax, mo.value)
        self._cell_sizes = self._idxs = None

    def compute(self, points, stream=None) -> None:
        """cellListGPU::compute(points) (cellListGPU.cpp:289-330): CPU tensors
        are binned on the host (same buckets, index order inside each), device
        tensors by the device build."""
        if isinstance(points, torch.Tensor) and points.is_cuda:
            self.computeGPU(points, stream)
            return
        L = _lib.load()
        n = int(points.shape[0])
        box = self.Box.box_matrix()
        bp = _box_arg(box)
        ptr = _host_ptr(points, torch.float64, "points")
        xs, ys, mo = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(L.gdt_host_celllist_general(ptr, n, bp, self.a, None, None, 0,
                                               ctypes.byref(xs), ctypes.byref(ys),
                                               ctypes.byref(mo)), "gdt_host_celllist")
        self.Np = n
        self.xsize, self.ysize = xs.value, ys.value
        self.boxsize = box[0] / xs.value if self.Box.is_rectangle() else 0.0
        self.Nmax = max(self.Nmax, mo.value)
        nc, nmax = self.xsize * self.ysize, max(1, self.Nmax)
        self._cell_sizes = torch.zeros(nc, dtype=torch.int32)
        self._idxs = torch.zeros(nc * nmax, dtype=torch.int32)
        _lib.check(L.gdt_host_celllist_general(ptr, n, bp, self.a, self._cell_sizes.data_ptr(),
                                               self._idxs.data_ptr(), nmax, ctypes.byref(xs),
                                               ctypes.byref(ys), ctypes.byref(mo)),
                   "gdt_host_celllist")

    def _export(self):
        L = _lib.load()
        dev = torch.device("cuda", self.ctx.device)
        nc = self.xsize * self.ysize
        nmax = max(1, self.Nmax)
        self._cell_sizes = torch.zeros(nc, dtype=torch.int32, device=dev)
        self._idxs = torch.zeros(nc * nmax, dtype=torch.int32, device=dev)
        _lib.check(L.gdt_celllist_export(self.ctx.handle, self._cell_sizes.data_ptr(),
                                         self._idxs.data_ptr(), nmax, _stream_ptr()),
                   "gdt_celllist_export")

    @property
    def cell_sizes(self):
        """Points per cell, cells row-major (cell_indexer(cx,cy) = cy*xsize+cx)."""
        if self._cell_sizes is None:
            self._export()
        return self._cell_sizes

    @property
    def idxs(self):
        """Padded bucket contents: idxs[cell*Nmax + k] (cell_list_indexer)."""
        if self._idxs is None:
            self._export()
        return self._idxs

    def getCellSizeArray(self):
        return self.cell_sizes

    def getIdxArray(self):
        return self.idxs


class DelaunayGPU:
    """Localized periodic Delaunay triangulation (DelaunayGPU, DelaunayGPU.h:18-130)."""

    def __init__(self, N: int | None = None, maximumNeighborsGuess: int = 16,
                 cellSize: float = 1.0, bx: PeriodicBoundaries | None = None, device: int = 0):
        self._ctx = None               # the device context, made on first device use
        self.device = int(device)
        self.prof = MultiProfiler()
        self.safetyMode = False
        self.GPUcompute = True
        self.OMPThreadsNum = 1
        self.precision = _lib.PREC_FP64
        self.Ncells = 0
        self.MaxSize = 4
        self.cellsize = 1.10           # reference default (DelaunayGPU.cpp:7)
        self.Box = PeriodicBoundaries(1.0)
        self.cList = CellListGPU(self.Box, lambda: self.ctx)
        self.last_repaired = 0
        if N is not None:
            self.initialize(N, maximumNeighborsGuess, cellSize, bx)

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(self.device)
        return self._ctx

    # --- set-up (DelaunayGPU.cpp:16-62) -----------------------------------
    def initialize(self, N: int, maximumNeighborsGuess: int, cellSize: float,
                   bx: PeriodicBoundaries | None) -> None:
        self.prof.start("initialization")
        self.Ncells = int(N)
        self.MaxSize = max(4, int(maximumNeighborsGuess))
        self.cellsize = float(cellSize)
        if bx is not None:
            self.setBox(bx)
        self._init_cell_list()
        self.prof.end("initialization")

    def _init_cell_list(self) -> None:
        self.cList.setNp(self.Ncells)
        self.cList.setBox(self.Box)
        self.cList.setGridSize(self.cellsize)

    def resize(self, nmax: int) -> None:
        self.MaxSize = int(nmax)

    def setCellListSize(self, csize: float) -> None:
        self.cellsize = float(csize)
        self._init_cell_list()

    def updateList(self, points) -> None:
        if self.GPUcompute:
            self.cList.computeGPU(points)
        else:
            self.cList.compute(points)

    def setBox(self, bx: PeriodicBoundaries) -> None:
        self.Box = bx
        self.cList.setBox(bx)

    def setSafetyMode(self, sm: bool) -> None:
        self.safetyMode = bool(sm)

    def setGPUcompute(self, flag: bool) -> None:
        """False: CPU mode, on CPU tensors, over setOMPthreads threads."""
        self.GPUcompute = bool(flag)

    def setOMPthreads(self, num: int) -> None:
        self.OMPThreadsNum = int(num)

    def setPrecision(self, precision: int) -> None:
        """64: all-FP64 predicates; 32: FP32 with certified bound + FP64
        fallback; 128 (PREC_EXACT): every decision exact -- FP64 where its
        error bound certifies it, expansion arithmetic otherwise, exact ties
        by the symbolic perturbation (rectangular boxes, coordinates within
        four box lengths of the box)."""
        if precision not in (_lib.PREC_FP64, _lib.PREC_FP32F, _lib.PREC_EXACT):
            raise ValueError("precision must be 64, 32 or 128")
        self.precision = int(precision)

    def stats(self) -> dict:
        return self.ctx.stats()

    # --- main entry points ------------------------------------------------
    def _check_args(self, points, tri, num):
        n = int(points.shape[0])
        if points.dim() != 2 or points.shape[1] != 2:
            raise ValueError("points must be [N, 2]")
        _dev_ptr(points, torch.float64, "points")
        _dev_ptr(tri, torch.int32, "GPUTriangulation")
        _dev_ptr(num, torch.int32, "cellNeighborNum")
        if num.numel() != n:
            raise ValueError("cellNeighborNum must have N entries")
        return n

    def _check_host_args(self, points, tri, num):
        n = int(points.shape[0])
        if points.dim() != 2 or points.shape[1] != 2:
            raise ValueError("points must be [N, 2]")
        _host_ptr(points, torch.float64, "points")
        _host_ptr(tri, torch.int32, "GPUTriangulation")
        _host_ptr(num, torch.int32, "cellNeighborNum")
        if num.numel() != n:
            raise ValueError("cellNeighborNum must have N entries")
        return n

    def globalDelaunayTriangulation(self, points, GPUTriangulation, cellNeighborNum,
                                    stream=None) -> None:
        """DelaunayGPU::globalDelaunayTriangulation (DelaunayGPU.cpp:118-157)."""
        n = int(points.shape[0])
        if n == 0:
            print("No points in GPU DT")
            return
        if n != self.Ncells or GPUTriangulation.numel() != n * self.MaxSize:
            self.Ncells = n
            GPUTriangulation.resize_(n * self.MaxSize)
            self._init_cell_list()
        self.prof.start("globalTriangulation")
        if self.GPUcompute:
            self._check_args(points, GPUTriangulation, cellNeighborNum)
        else:
            self._check_host_args(points, GPUTriangulation, cellNeighborNum)
        L = _lib.load()
        x11, _, _, x22 = self.Box.getBoxDims()
        rect = self.Box.is_rectangle()
        box, off = _box_arg(self.Box.box_matrix()), (ctypes.c_int * 2)(0, n)
        while True:
            req = ctypes.c_int(0)
            if self.GPUcompute and rect:
                st = L.gdt_triangulate(self.ctx.handle, points.data_ptr(), n, x11, x22,
                                       self.cellsize, self.MaxSize, GPUTriangulation.data_ptr(),
                                       cellNeighborNum.data_ptr(), self.precision,
                                       ctypes.byref(req), _stream_ptr(stream))
            elif self.GPUcompute:
                st = L.gdt_triangulate_general(self.ctx.handle, points.data_ptr(), 1, off, box,
                                               self.cellsize, self.MaxSize,
                                               GPUTriangulation.data_ptr(),
                                               cellNeighborNum.data_ptr(), self.precision,
                                               ctypes.byref(req), _stream_ptr(stream))
            else:
                st = L.gdt_host_triangulate_general(points.data_ptr(), n, box, self.cellsize,
                                                    self.MaxSize, GPUTriangulation.data_ptr(),
                                                    cellNeighborNum.data_ptr(), self.precision,
                                                    self.OMPThreadsNum, ctypes.byref(req))
            if st == _lib.GDT_OVERFLOW:
                self.resize(req.value)
                GPUTriangulation.resize_(n * self.MaxSize)
                continue
            self.prof.end("globalTriangulation")
            _lib.check(st, "globalDelaunayTriangulation")
            return

    def testAndRepairDelaunayTriangulation(self, points, GPUTriangulation, cellNeighborNum,
                                           stream=None) -> None:
        """DelaunayGPU::testAndRepairDelaunayTriangulation (DelaunayGPU.cpp:567-602)."""
        if self.GPUcompute:
            n = self._check_args(points, GPUTriangulation, cellNeighborNum)
        else:
            n = self._check_host_args(points, GPUTriangulation, cellNeighborNum)
        if GPUTriangulation.numel() != n * self.MaxSize:
            raise ValueError("GPUTriangulation must hold N*MaxSize entries")
        L = _lib.load()
        x11, _, _, x22 = self.Box.getBoxDims()
        box, off = _box_arg(self.Box.box_matrix()), (ctypes.c_int * 2)(0, n)
        nf, req = ctypes.c_int(0), ctypes.c_int(0)
        self.prof.start("testAndRepair")
        if self.GPUcompute and self.Box.is_rectangle():
            st = L.gdt_test_and_repair(self.ctx.handle, points.data_ptr(), n, x11, x22,
                                       self.cellsize, self.MaxSize, GPUTriangulation.data_ptr(),
                                       cellNeighborNum.data_ptr(), self.precision,
                                       ctypes.byref(nf), ctypes.byref(req), _stream_ptr(stream))
        elif self.GPUcompute:
            st = L.gdt_test_and_repair_general(self.ctx.handle, points.data_ptr(), 1, off, box,
                                               self.cellsize, self.MaxSize,
                                               GPUTriangulation.data_ptr(),
                                               cellNeighborNum.data_ptr(), self.precision,
                                               ctypes.byref(nf), ctypes.byref(req),
                                               _stream_ptr(stream))
        else:
            st = L.gdt_host_test_and_repair_general(points.data_ptr(), n, box, self.cellsize,
                                                    self.MaxSize, GPUTriangulation.data_ptr(),
                                                    cellNeighborNum.data_ptr(), self.precision,
                                                    self.OMPThreadsNum, ctypes.byref(nf),
                                                    ctypes.byref(req))
        self.prof.end("testAndRepair")
        self.last_repaired = nf.value
        if st == _lib.GDT_OVERFLOW:
            # a repaired ring outgrew the row stride: re-triangulate globally
            self.resize(req.value)
            self.globalDelaunayTriangulation(points, GPUTriangulation, cellNeighborNum, stream)
            return
        _lib.check(st, "testAndRepairDelaunayTriangulation")


def _box_tuple(b) -> tuple:
    """A batch box as (lx, ly) when it is a rectangle, else (x11, x12, x21, x22)."""
    if isinstance(b, PeriodicBoundaries):
        b = b.getBoxDims()
    b = tuple(float(v) for v in b)
    if len(b) == 4 and b[1] == 0.0 and b[2] == 0.0 and b[0] > 0 and b[3] > 0:
        return (b[0], b[3])
    if len(b) not in (2, 4):
        raise ValueError("a box is (lx, ly) or (x11, x12, x21, x22)")
    return b


def _entry_point(name: str, general: bool) -> str:
    """The C entry point of a batched call for rectangles or general boxes."""
    if general:
        return name + "_general"
    return name if name == "gdt_md_run" else name + "_batch"


MD_HEADROOM = 4   # row slots beyond the longest ring, given before and after the Brownian loop runs


class BatchedDelaunayGPU:
    """Many independent periodic sets in one launch per kernel (no reference
    counterpart: the reference runs one DelaunayGPU per set).

    Set s is rows ``offsets[s]:offsets[s+1]`` of the concatenated points, in
    the box ``boxes[s]``: ``(lx, ly)``, a box matrix ``(x11, x12, x21, x22)``
    (lattice vectors as columns) or a PeriodicBoundaries; ring entries index
    the concatenated array.
    Same triangulation / test-and-repair semantics as DelaunayGPU, including
    the growth of ``MaxSize`` (the row stride shared by every set)."""

    def __init__(self, sizes, boxes, maximumNeighborsGuess: int = 16, cellSize: float = 1.0,
                 device: int = 0, precision: int = _lib.PREC_FP64):
        sizes = [int(s) for s in sizes]
        if len(sizes) != len(boxes) or not sizes:
            raise ValueError("one box per set")
        self.offsets = [0]
        for s in sizes:
            self.offsets.append(self.offsets[-1] + s)
        self.boxes = [_box_tuple(b) for b in boxes]
        self.N = self.offsets[-1]
        self.nsets = len(sizes)
        self.MaxSize = max(4, int(maximumNeighborsGuess))
        self.cellsize = float(cellSize)
        self.ctx = Context(device)
        self.precision = int(precision)
        self.last_repaired = 0
        self._off = (ctypes.c_int * len(self.offsets))(*self.offsets)
        # every box a rectangle: (lx, ly) pairs and the rectangular calls;
        # else box matrices (4 doubles per set) and the general calls
        self.general = not all(len(b) == 2 for b in self.boxes)
        mats = [(b[0], 0.0, 0.0, b[1]) if len(b) == 2 else b for b in self.boxes]
        flat = [v for b in (mats if self.general else self.boxes) for v in b]
        self._box = (ctypes.c_double * len(flat))(*flat)
        self._tri = _entry_point("gdt_triangulate", self.general)
        self._tar = _entry_point("gdt_test_and_repair", self.general)
        self._md = _entry_point("gdt_md_run", self.general)

    def setPrecision(self, precision: int) -> None:
        if precision not in (_lib.PREC_FP64, _lib.PREC_FP32F, _lib.PREC_EXACT):
            raise ValueError("precision must be 64, 32 or 128")
        self.precision = int(precision)

    def stats(self) -> dict:
        return self.ctx.stats()

    def _check(self, points, tri, num):
        if points.dim() != 2 or points.shape[0] != self.N or points.shape[1] != 2:
            raise ValueError(f"points must be [{self.N}, 2]")
        _dev_ptr(points, torch.float64, "points")
        _dev_ptr(tri, torch.int32, "GPUTriangulation")
        _dev_ptr(num, torch.int32, "cellNeighborNum")
        if num.numel() != self.N:
            raise ValueError("cellNeighborNum must have N entries")

    def globalDelaunayTriangulation(self, points, GPUTriangulation, cellNeighborNum,
                                    stream=None) -> None:
        if GPUTriangulation.numel() != self.N * self.MaxSize:
            GPUTriangulation.resize_(self.N * self.MaxSize)
        self._check(points, GPUTriangulation, cellNeighborNum)
        L = _lib.load()
        while True:
            req = ctypes.c_int(0)
            st = getattr(L, self._tri)(self.ctx.handle, points.data_ptr(), self.nsets, self._off,
                                         self._box, self.cellsize, self.MaxSize,
                                         GPUTriangulation.data_ptr(), cellNeighborNum.data_ptr(),
                                         self.precision, ctypes.byref(req), _stream_ptr(stream))
            if st == _lib.GDT_OVERFLOW:
                self.MaxSize = int(req.value)
                GPUTriangulation.resize_(self.N * self.MaxSize)
                continue
            _lib.check(st, "globalDelaunayTriangulation (batch)")
            return

    def testAndRepairDelaunayTriangulation(self, points, GPUTriangulation, cellNeighborNum,
                                           stream=None) -> None:
        self._check(points, GPUTriangulation, cellNeighborNum)
        if GPUTriangulation.numel() != self.N * self.MaxSize:
            raise ValueError("GPUTriangulation must hold N*MaxSize entries")
        L = _lib.load()
        nf, req = ctypes.c_int(0), ctypes.c_int(0)
        st = getattr(L, self._tar)(self.ctx.handle, points.data_ptr(), self.nsets, self._off,
                                         self._box, self.cellsize, self.MaxSize,
                                         GPUTriangulation.data_ptr(), cellNeighborNum.data_ptr(),
                                         self.precision, ctypes.byref(nf), ctypes.byref(req),
                                         _stream_ptr(stream))
        self.last_repaired = nf.value
        if st == _lib.GDT_OVERFLOW:
            self.MaxSize = int(req.value)
            self.globalDelaunayTriangulation(points, GPUTriangulation, cellNeighborNum, stream)
            return
        _lib.check(st, "testAndRepairDelaunayTriangulation (batch)")

    def restride(self, GPUTriangulation, new_max: int) -> None:
        """Widen the rows of a valid ring table to ``new_max`` slots in
        place (row i moves to ``i*new_max``), as the reference's resize of
        its triangulation array does (DelaunayGPU.cpp:111, 130)."""
        new_max = int(new_max)
        if new_max <= self.MaxSize:
            return
        rows = GPUTriangulation.view(self.N, self.MaxSize).clone()
        GPUTriangulation.resize_(self.N * new_max)
        GPUTriangulation.view(self.N, new_max)[:, :self.MaxSize].copy_(rows)
        self.MaxSize = new_max

    # --- Brownian test-and-repair loop -----------------------------------
    def md_run(self, points, GPUTriangulation, cellNeighborNum, nsteps: int, sigma: float = 0.01,
               seed: int = 1, step0: int = 0, use_graph: bool = True, check_every: int = 100,
               stream=None, headroom: int = MD_HEADROOM) -> dict:
        """``nsteps`` steps of: move every point by sigma*N(0,1) per coordinate
        (counter-based noise keyed by (seed, step, set, point)), then
        test-and-repair, all on the device (one hipGraph replay per step when
        ``use_graph``).  The rings must triangulate ``points`` on entry.

        Before the loop the rows are widened to the longest ring plus
        ``headroom`` slots (0: as they are).  The first step that fails (a
        repaired ring longer than ``MaxSize``, an error) halts the loop on
        the device at its end: no later step runs, the positions keep that
        step's displacement, and the host completes the step with a global
        triangulation of the current positions (after widening the rows by
        ``headroom`` beyond the longest ring on an overflow), then goes on
        from the next step.  Returns the totals (steps, repaired points,
        recoveries, and for each recovery the failing step)."""
        self._check(points, GPUTriangulation, cellNeighborNum)
        if GPUTriangulation.numel() != self.N * self.MaxSize:
            raise ValueError("GPUTriangulation must hold N*MaxSize entries")
        L = _lib.load()
        if headroom > 0:
            self.restride(GPUTriangulation, int(cellNeighborNum.max().item()) + int(headroom))
        out = dict(steps=0, flagged_total=0, max_flagged=0, recoveries=0, recovered=[],
                   skipped_replays=0)
        step, left = int(step0), int(nsteps)
        while left > 0:
            rep = _lib.MdReport()
            st = getattr(L, self._md)(self.ctx.handle, points.data_ptr(), self.nsets, self._off,
                                      self._box,
                              self.cellsize, self.MaxSize, GPUTriangulation.data_ptr(),
                              cellNeighborNum.data_ptr(), self.precision, float(sigma),
                              int(seed) & 0xFFFFFFFFFFFFFFFF, step, left, int(bool(use_graph)),
                              int(check_every), ctypes.byref(rep), _stream_ptr(stream))
            step += rep.steps_done
            left -= rep.steps_done
            out["steps"] += rep.steps_done
            out["flagged_total"] += rep.flagged_total
            out["max_flagged"] = max(out["max_flagged"], rep.max_flagged)
            out["skipped_replays"] += rep.skipped_steps
            if st == _lib.GDT_OK:
                break
            if st not in (_lib.GDT_OVERFLOW, _lib.GDT_ERR_NONGENERIC, _lib.GDT_ERR_EULER):
                _lib.check(st, "md_run")
            if rep.first_bad_step != step - 1:
                raise RuntimeError(f"md_run: the loop halted at step {rep.first_bad_step} but "
                                   f"the positions are at step {step}")
            if st == _lib.GDT_OVERFLOW:
                # rings keep changing as the points move: leave headroom (the
                # truncated rows are rebuilt below, so no row needs moving)
                self.MaxSize = max(self.MaxSize, int(rep.required)) + max(int(headroom), 1)
                GPUTriangulation.resize_(self.N * self.MaxSize)
            out["recoveries"] += 1
            out["recovered"].append(dict(status=_lib.status_string(st), failed_step=
                                         rep.first_bad_step, required=rep.required,
                                         max_size=self.MaxSize))
            # completes the failing step: the triangulation of its positions
            self.globalDelaunayTriangulation(points, GPUTriangulation, cellNeighborNum, stream)
        self.last_repaired = out["max_flagged"]
        return out
