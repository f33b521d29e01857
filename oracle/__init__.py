"""ctypes wrapper of the CPU oracle (oracle/libgdt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The shipped library never
imports it.  Functions mirror gdt_oracle.h; each wraps the reference
function named there.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgdt_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_triangulate.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int]
        L.orc_triangulate.restype = ctypes.c_int
        L.orc_triangles.argtypes = [_ip, _ip, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_longlong)]
        L.orc_triangles.restype = ctypes.c_int
        L.orc_test_circumcircles.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, _ip, ctypes.c_int, _ip, ctypes.c_int]
        L.orc_test_circumcircles.restype = ctypes.c_int
        L.orc_repair.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _ip, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int]
        L.orc_repair.restype = ctypes.c_int
        L.orc_test_and_repair.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _ip, _ip, _ip, _ip,
                                          ctypes.c_int]
        L.orc_test_and_repair.restype = ctypes.c_int
        L.orc_voronoi_energy.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         _ip, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         _dp, _dp, _dp, ctypes.c_int]
        L.orc_voronoi_energy.restype = ctypes.c_double
        L.orc_triangulate_box.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                          _ip, _ip, _ip, ctypes.c_int]
        L.orc_triangulate_box.restype = ctypes.c_int
        L.orc_test_circumcircles_box.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double, _ip,
                                                 ctypes.c_int, _ip, ctypes.c_int]
        L.orc_test_circumcircles_box.restype = ctypes.c_int
        L.orc_test_and_repair_box.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double,
                                              ctypes.c_int, _ip, _ip, _ip, _ip, ctypes.c_int]
        L.orc_test_and_repair_box.restype = ctypes.c_int
        L.orc_voronoi_energy_box.argtypes = [_dp, ctypes.c_int, _dp, _ip, _ip, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                             ctypes.c_int]
        L.orc_voronoi_energy_box.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def triangulate(xy, lx, ly=None, cell_size=1.0, max_nbrs=16, nthreads=0):
    """Global triangulation; grows max_nbrs and retries on overflow like the
    reference's safetyMode loop (DelaunayGPU.cpp:137-156).
    Returns (nbrs [n, max_nbrs] int32, counts [n] int32)."""
    ly = lx if ly is None else ly
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    m = max(4, int(max_nbrs))
    while True:
        nbrs = np.full((n, m), -1, dtype=np.int32)
        counts = np.zeros(n, dtype=np.int32)
        req = np.zeros(1, dtype=np.int32)
        rc = lib().orc_triangulate(_d(xy), n, lx, ly, cell_size, m, _i(nbrs), _i(counts),
                                   _i(req), nthreads)
        if rc == 1:
            m = int(req[0])
            continue
        if rc != 0:
            raise RuntimeError(f"oracle triangulate failed ({rc})")
        return nbrs, counts


def triangles(nbrs, counts):
    """Owned-triangle list (getCircumcirclesCPU) and sum(counts)."""
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    n, stride = nbrs.shape
    cap = int(counts.sum()) // 3 + 8
    tris = np.zeros((cap, 3), dtype=np.int32)
    tot = ctypes.c_longlong(0)
    t = lib().orc_triangles(_i(nbrs), _i(counts), n, stride, _i(tris), cap, ctypes.byref(tot))
    if t < 0:
        raise RuntimeError("triangle buffer too small")
    return tris[:t].copy(), int(tot.value)


def test_circumcircles(xy, lx, ly, tris, cell_size=1.0, nthreads=0):
    """repair flags (i or -1) from the empty-circle test of every triangle."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    rep = np.zeros(xy.shape[0], dtype=np.int32)
    lib().orc_test_circumcircles(_d(xy), xy.shape[0], lx, ly, cell_size, _i(tris), tris.shape[0],
                                 _i(rep), nthreads)
    return rep


def test_and_repair(xy, lx, ly, nbrs, counts, cell_size=1.0, nthreads=0):
    """In-place testAndRepair on copies; returns (nbrs, counts, nflagged)."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32).copy()
    counts = np.ascontiguousarray(counts, dtype=np.int32).copy()
    n, m = nbrs.shape
    req = np.zeros(1, dtype=np.int32)
    nf = np.zeros(1, dtype=np.int32)
    rc = lib().orc_test_and_repair(_d(xy), n, lx, ly, cell_size, m, _i(nbrs), _i(counts),
                                   _i(req), _i(nf), nthreads)
    if rc == -2:
        raise RuntimeError("Euler check failed")
    if rc == 1:
        raise RuntimeError(f"repair needs max_nbrs={int(req[0])}")
    if rc != 0:
        raise RuntimeError(f"oracle test_and_repair failed ({rc})")
    return nbrs, counts, int(nf[0])


def voronoi_energy(xy, lx, ly, nbrs, counts, a0=1.0, p0=3.8, nthreads=0):
    """(E, area, perim, grad) with the topology given by the rings."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    n, stride = nbrs.shape
    area = np.zeros(n)
    perim = np.zeros(n)
    grad = np.zeros((n, 2))
    E = lib().orc_voronoi_energy(_d(xy), n, lx, ly, _i(nbrs), _i(counts), stride, a0, p0,
                                 _d(area), _d(perim), _d(grad), nthreads)
    return E, area, perim, grad


# ---- general (triclinic) boxes: box = (x11, x12, x21, x22), lattice vectors
# as columns (periodicBoundaries::setGeneral, periodicBoundaries.h:86-101) ----

def _box(box):
    b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(4))
    return b


def triangulate_box(xy, box, cell_size=1.0, max_nbrs=16, nthreads=0):
    """triangulate() in a general box."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    b = _box(box)
    n = xy.shape[0]
    m = max(4, int(max_nbrs))
    while True:
        nbrs = np.full((n, m), -1, dtype=np.int32)
        counts = np.zeros(n, dtype=np.int32)
        req = np.zeros(1, dtype=np.int32)
        rc = lib().orc_triangulate_box(_d(xy), n, _d(b), cell_size, m, _i(nbrs), _i(counts),
                                       _i(req), nthreads)
        if rc == 1:
            m = int(req[0])
            continue
        if rc != 0:
            raise RuntimeError(f"oracle triangulate failed ({rc})")
        return nbrs, counts


def test_circumcircles_box(xy, box, tris, cell_size=1.0, nthreads=0):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    b = _box(box)
    rep = np.zeros(xy.shape[0], dtype=np.int32)
    lib().orc_test_circumcircles_box(_d(xy), xy.shape[0], _d(b), cell_size, _i(tris),
                                     tris.shape[0], _i(rep), nthreads)
    return rep


def test_and_repair_box(xy, box, nbrs, counts, cell_size=1.0, nthreads=0):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32).copy()
    counts = np.ascontiguousarray(counts, dtype=np.int32).copy()
    b = _box(box)
    n, m = nbrs.shape
    req = np.zeros(1, dtype=np.int32)
    nf = np.zeros(1, dtype=np.int32)
    rc = lib().orc_test_and_repair_box(_d(xy), n, _d(b), cell_size, m, _i(nbrs), _i(counts),
                                       _i(req), _i(nf), nthreads)
    if rc == -2:
        raise RuntimeError("Euler check failed")
    if rc == 1:
        raise RuntimeError(f"repair needs max_nbrs={int(req[0])}")
    if rc != 0:
        raise RuntimeError(f"oracle test_and_repair failed ({rc})")
    return nbrs, counts, int(nf[0])


def voronoi_energy_box(xy, box, nbrs, counts, a0=1.0, p0=3.8, nthreads=0):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    b = _box(box)
    n, stride = nbrs.shape
    area = np.zeros(n)
    perim = np.zeros(n)
    grad = np.zeros((n, 2))
    E = lib().orc_voronoi_energy_box(_d(xy), n, _d(b), _i(nbrs), _i(counts), stride, a0, p0,
                                     _d(area), _d(perim), _d(grad), nthreads)
    return E, area, perim, grad
