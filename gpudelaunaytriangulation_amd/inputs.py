"""Deterministic synthetic point sets for the triangulation workloads.

Every generator is a pure function of (seed, index) built from SplitMix64 and
IEEE basic operations (+, -, *, /, sqrt, frexp/ldexp), so the same call gives
the same bits on any x86-64 host.  The Gaussian jitter uses the Marsaglia
polar method with a logarithm evaluated here from its atanh series instead of
the C library's ``log`` (whose last bit may depend on the CPU's code path).

Workloads (box side L, density 1 unless noted):

* ``uniform_points(n, lx, ly, seed)``: i.i.d. uniform in [0,lx) x [0,ly).
* ``lattice_points(nx, ny, sigma, seed)``: triangular lattice of spacing 1,
  rows sqrt(3)/2 apart, odd rows shifted by 1/2, per-coordinate Gaussian
  jitter, wrapped into the box nx x ny*sqrt(3)/2.
* ``batch_sizes(nsets, nmin, nmax, seed)``: set sizes log-uniform in
  [nmin, nmax].
"""
from __future__ import annotations

import hashlib

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_LN2_HI = 6.93147180369123816490e-01
_LN2_LO = 1.90821492927058770002e-10


def splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """SplitMix64 output for state seed + (counter+1)*golden (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counters.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, counters: np.ndarray) -> np.ndarray:
    """53-bit uniforms in [0,1): (r >> 11) * 2^-53 (exact)."""
    r = splitmix64(seed, counters) >> np.uint64(11)
    return r.astype(np.float64) * (1.0 / 9007199254740992.0)


def det_log(x: np.ndarray) -> np.ndarray:
    """Natural log from basic IEEE operations only (x > 0).

    x = m * 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(t),
    t = (m-1)/(m+1), |t| <= 0.1716, summed to t^29 (truncation < 1e-23).
    """
    m, e = np.frexp(np.asarray(x, dtype=np.float64))
    small = m < 0.7071067811865476
    m = np.where(small, m * 2.0, m)
    e = (e - small.astype(e.dtype)).astype(np.float64)
    t = (m - 1.0) / (m + 1.0)
    t2 = t * t
    s = np.zeros_like(t)
    for k in range(29, 1, -2):
        s = (s + 1.0 / k) * t2
    s = (s + 1.0) * t
    return e * _LN2_HI + (e * _LN2_LO + 2.0 * s)


def gaussian(seed: int, n: int) -> np.ndarray:
    """n standard normals (Marsaglia polar method, counter-indexed rejection)."""
    out = np.empty(n, dtype=np.float64)
    filled = 0
    base = 0
    while filled < n:
        need = n - filled
        batch = max(1024, int(need * 0.7) + 64)
        c = np.arange(base, base + batch, dtype=np.uint64)
        base += batch
        v1 = 2.0 * uniform01(seed, 2 * c) - 1.0
        v2 = 2.0 * uniform01(seed, 2 * c + 1) - 1.0
        s = v1 * v1 + v2 * v2
        ok = (s > 0.0) & (s < 1.0)
        v1, v2, s = v1[ok], v2[ok], s[ok]
        f = np.sqrt(-2.0 * det_log(s) / s)
        g = np.empty(2 * s.size, dtype=np.float64)
        g[0::2] = v1 * f
        g[1::2] = v2 * f
        take = min(need, g.size)
        out[filled:filled + take] = g[:take]
        filled += take
    return out


def wrap_into_box(xy: np.ndarray, lx: float, ly: float) -> np.ndarray:
    """Wrap coordinates into [0,lx) x [0,ly)."""
    out = xy.copy()
    for col, L in ((0, lx), (1, ly)):
        v = out[:, col]
        v = v - np.floor(v / L) * L
        v = np.where(v >= L, v - L, v)
        v = np.where(v < 0.0, 0.0, v)
        out[:, col] = v
    return out


def uniform_points(n: int, lx: float, ly: float | None = None, seed: int = 1) -> np.ndarray:
    """n points i.i.d. uniform in [0,lx) x [0,ly), float64, shape (n,2)."""
    ly = lx if ly is None else ly
    c = np.arange(2 * n, dtype=np.uint64)
    u = uniform01(seed, c).reshape(n, 2)
    xy = np.empty((n, 2), dtype=np.float64)
    xy[:, 0] = u[:, 0] * lx
    xy[:, 1] = u[:, 1] * ly
    return wrap_into_box(xy, lx, ly)


LATTICE_ROW = 0.8660254037844386  # sqrt(3)/2


def lattice_box(nx: int, ny: int) -> tuple[float, float]:
    """Periodic box of an nx x ny triangular lattice with unit spacing."""
    return float(nx), float(ny) * LATTICE_ROW


def lattice_points(nx: int, ny: int, sigma: float = 1e-5, seed: int = 3) -> np.ndarray:
    """Jittered triangular lattice (ny even for periodicity), shape (nx*ny, 2)."""
    r, c = np.divmod(np.arange(nx * ny, dtype=np.int64), nx)
    g = gaussian(seed, 2 * nx * ny).reshape(-1, 2)
    xy = np.empty((nx * ny, 2), dtype=np.float64)
    xy[:, 0] = c.astype(np.float64) + 0.5 * (r % 2).astype(np.float64) + sigma * g[:, 0]
    xy[:, 1] = r.astype(np.float64) * LATTICE_ROW + sigma * g[:, 1]
    lx, ly = lattice_box(nx, ny)
    return wrap_into_box(xy, lx, ly)


def batch_sizes(nsets: int, nmin: int = 1024, nmax: int = 65536, seed: int = 4) -> np.ndarray:
    """Set sizes log-uniform in [nmin, nmax] (rounded to integers)."""
    u = uniform01(seed, np.arange(nsets, dtype=np.uint64))
    sizes = np.floor(nmin * np.exp(u * np.log(nmax / nmin)) + 0.5).astype(np.int64)
    return np.clip(sizes, nmin, nmax)


def batch_points(sizes: np.ndarray, seed: int = 4) -> tuple[list[np.ndarray], list[float]]:
    """Uniform points per set in a square box of side sqrt(N_s)."""
    pts, boxes = [], []
    for s, n in enumerate(sizes):
        L = float(np.sqrt(float(n)))
        pts.append(uniform_points(int(n), L, L, seed=seed * 1000003 + s))
        boxes.append(L)
    return pts, boxes


MD_ATTEMPTS = 64


def md_key(seed: int, step: int, set_index: int) -> int:
    """Stream key of (seed, step, set) for the Brownian loop (gdt_rng.h)."""
    c = np.array([(int(step) << 20) | int(set_index)], dtype=np.uint64)
    return int(splitmix64(seed, c)[0])


def md_noise(seed: int, step: int, set_index: int, n: int) -> np.ndarray:
    """[n, 2] standard normals of one set at one step of the Brownian loop:
    point i takes the first accepted polar-method attempt a < 64 over the
    counters 2*(i*64 + a) and 2*(i*64 + a) + 1 of its stream; none accepted
    gives (0, 0).  Bit-identical to gdt_gauss_pair (gdt_rng.h)."""
    key = md_key(seed, step, set_index)
    g = np.zeros((n, 2), dtype=np.float64)
    todo = np.arange(n, dtype=np.uint64)
    for a in range(MD_ATTEMPTS):
        if todo.size == 0:
            break
        c = todo * np.uint64(MD_ATTEMPTS) + np.uint64(a)
        v1 = 2.0 * uniform01(key, np.uint64(2) * c) - 1.0
        v2 = 2.0 * uniform01(key, np.uint64(2) * c + np.uint64(1)) - 1.0
        s = v1 * v1 + v2 * v2
        ok = (s > 0.0) & (s < 1.0)
        so = s[ok]
        f = np.sqrt(-2.0 * det_log(so) / so)
        g[todo[ok], 0] = v1[ok] * f
        g[todo[ok], 1] = v2[ok] * f
        todo = todo[~ok]
    return g


def md_step(xy: np.ndarray, lx: float, ly: float, sigma: float, seed: int, step: int,
            set_index: int) -> np.ndarray:
    """One Brownian displacement of one set, wrapped into the box."""
    return wrap_into_box(xy + sigma * md_noise(seed, step, set_index, xy.shape[0]), lx, ly)


def box_inverse(box) -> tuple:
    """Inverse of a box matrix (x11, x12, x21, x22), in the library's operation
    order (gdt_make_set_box)."""
    b0, b1, b2, b3 = (float(v) for v in box)
    bd = b0 * b3 - b1 * b2
    return (b3 / bd, -b1 / bd, -b2 / bd, b0 / bd)


def general_points(n: int, box, seed: int = 1) -> np.ndarray:
    """n points uniform in the unit cell of the general box (x11, x12, x21, x22)
    (lattice vectors as columns): uniform unit-cell coordinates mapped by the
    box matrix."""
    X = np.asarray(box, dtype=np.float64).reshape(2, 2)
    return uniform_points(n, 1.0, 1.0, seed) @ X.T


def wrap_into_general_box(xy: np.ndarray, box) -> np.ndarray:
    """Wrap points into the unit cell of a general box by whole lattice vectors
    (bit-identical to gdt_wrap_into_box): k = floor(unit-cell coordinates),
    p -= k_u a + k_v b."""
    b0, b1, b2, b3 = (float(v) for v in box)
    i0, i1, i2, i3 = box_inverse(box)
    x, y = xy[:, 0].copy(), xy[:, 1].copy()
    ku = np.floor(i0 * x + i1 * y)
    kv = np.floor(i2 * x + i3 * y)
    x = x - (ku * b0 + kv * b1)
    y = y - (ku * b2 + kv * b3)
    return np.stack([x, y], 1)


def md_step_general(xy: np.ndarray, box, sigma: float, seed: int, step: int,
                    set_index: int) -> np.ndarray:
    """md_step in a general box."""
    return wrap_into_general_box(xy + sigma * md_noise(seed, step, set_index, xy.shape[0]), box)


def sha256(a: np.ndarray) -> str:
    """SHA-256 of an array's bytes in C order."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
