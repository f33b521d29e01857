"""Ring utilities: canonical rotation, hashing and invariant checks.

Rings are stored as the reference stores them (DelaunayGPU.cpp:47,
indexer.h:38-41): a flat int32 array read as [N][stride], neighbour m of
point i at i*stride+m, with counts[i] valid entries in counter-clockwise
order from an unspecified start.  The canonical form rotates every ring to
start at its smallest neighbour index, which removes the start offset the way
the reference's own comparison does (triangulation.cpp:108-119).
"""
from __future__ import annotations

import hashlib

import numpy as np


def canonical(nbrs: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Flat int32 array: for each point in order, its ring rotated to start
    at the smallest index (concatenated, sum(counts) entries)."""
    nbrs = np.asarray(nbrs)
    counts = np.asarray(counts, dtype=np.int64)
    n, stride = nbrs.shape
    if n == 0:
        return np.zeros(0, dtype=np.int32)
    if counts.min() <= 0 or counts.max() > stride:
        raise ValueError("ring counts out of range")
    cmax = int(counts.max())
    m = np.arange(cmax)[None, :]
    valid = m < counts[:, None]
    big = np.iinfo(np.int32).max
    block = np.where(valid, nbrs[:, :cmax], big)
    start = np.argmin(block, axis=1)
    idx = (start[:, None] + m) % counts[:, None]
    rot = np.take_along_axis(nbrs[:, :cmax], idx, axis=1)
    return rot[valid].astype(np.int32)


def ring_hash(nbrs: np.ndarray, counts: np.ndarray) -> str:
    """SHA-256 over counts (int32) followed by the canonical rings (int32)."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(counts, dtype=np.int32).tobytes())
    h.update(canonical(nbrs, counts).tobytes())
    return h.hexdigest()


def rings_as_lists(nbrs: np.ndarray, counts: np.ndarray) -> list[list[int]]:
    return [list(map(int, nbrs[i, :counts[i]])) for i in range(len(counts))]


def check_torus_invariants(nbrs: np.ndarray, counts: np.ndarray) -> None:
    """Raise AssertionError unless the rings form a consistent triangulation
    of the torus: sum(counts) = 6N, symmetric adjacency, every triangle seen
    from its three vertices with the same orientation, 2N owned triangles."""
    nbrs = np.asarray(nbrs)
    counts = np.asarray(counts)
    n = len(counts)
    assert int(counts.sum()) == 6 * n, f"sum(counts)={int(counts.sum())} != 6N={6 * n}"
    succ = {}
    for i in range(n):
        r = nbrs[i, :counts[i]]
        assert len(set(r.tolist())) == len(r), f"ring {i} repeats a neighbour"
        for m in range(len(r)):
            succ[(i, int(r[m]))] = int(r[(m + 1) % len(r)])
    owned = 0
    for (i, a), b in succ.items():
        # triangle (i, a, b) CCW: seen from a it is (a, b, i), from b (b, i, a)
        assert succ.get((a, b)) == i, f"triangle ({i},{a},{b}) not seen from {a}"
        assert succ.get((b, i)) == a, f"triangle ({i},{a},{b}) not seen from {b}"
        if i < a and i < b:
            owned += 1
    assert owned == 2 * n, f"{owned} owned triangles != 2N={2 * n}"
