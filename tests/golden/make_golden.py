"""Write the golden fixtures under tests/golden/ (run in the build container).

    python tests/golden/make_golden.py

Every ring fixture is produced by the CPU oracle (oracle/, a restatement of
the reference's OpenMP path) and is written only after it has been checked
against an independent Delaunay triangulation from scipy's Qhull on a
periodically padded copy of the points and against the torus invariants.
The reference ships no golden vectors of its own (it has no tests); its
third-party cross-check (CGAL) is not installed here.

Fixtures:
  golden.json       point-array SHA-256 per workload, ring hashes, degree
                    histograms, and the C4 set sizes
  c1_n4096.npz      counts and canonical rings of the 4096-point workload
  c2_samples.npz    1024 sampled canonical rings of the 2^20-point workload
  c3b_samples.npz   1024 sampled canonical rings of the jittered lattice
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from gpudelaunaytriangulation_amd import inputs, rings  # noqa: E402
from qhull_periodic import periodic_rings  # noqa: E402


def workloads():
    lx3, ly3 = inputs.lattice_box(1024, 1024)
    return {
        "c1": dict(xy=lambda: inputs.uniform_points(4096, 64.0, 64.0, seed=1), lx=64.0, ly=64.0),
        "c2": dict(xy=lambda: inputs.uniform_points(1 << 20, 1024.0, 1024.0, seed=2),
                   lx=1024.0, ly=1024.0),
        "c3b": dict(xy=lambda: inputs.lattice_points(1024, 1024, 1e-5, seed=3), lx=lx3, ly=ly3),
    }


def sample_rings(nbrs, counts, k=1024, seed=11):
    n = counts.shape[0]
    idx = np.sort(np.random.default_rng(seed).choice(n, size=k, replace=False)).astype(np.int32)
    sub_n = nbrs[idx]
    sub_c = counts[idx]
    return idx, sub_c, rings.canonical(sub_n, sub_c)


def main():
    out = {"generator": "gpudelaunaytriangulation_amd.inputs (SplitMix64, polar Gaussian)",
           "oracle": "oracle/gdt_oracle.c, cell size 1.0", "workloads": {}}
    for name, w in workloads().items():
        t0 = time.time()
        xy = w["xy"]()
        lx, ly = w["lx"], w["ly"]
        nbrs, counts = oracle.triangulate(xy, lx, ly, 1.0, 16, nthreads=0)
        t1 = time.time()
        qn, qc = periodic_rings(xy, lx, ly)
        t2 = time.time()
        assert np.array_equal(counts, qc), f"{name}: oracle and Qhull degrees differ"
        assert np.array_equal(rings.canonical(nbrs, counts), rings.canonical(qn, qc)), \
            f"{name}: oracle and Qhull rings differ"
        assert int(counts.sum()) == 6 * xy.shape[0]
        if xy.shape[0] <= 4096:
            rings.check_torus_invariants(nbrs, counts)
        hist = np.bincount(counts).tolist()
        out["workloads"][name] = {
            "n": int(xy.shape[0]), "lx": lx, "ly": ly,
            "points_sha256": inputs.sha256(xy),
            "ring_sha256": rings.ring_hash(nbrs, counts),
            "degree_histogram": hist,
        }
        if name == "c1":
            np.savez_compressed(os.path.join(HERE, "c1_n4096.npz"), counts=counts,
                                canonical=rings.canonical(nbrs, counts))
        else:
            idx, sc, can = sample_rings(nbrs, counts)
            np.savez_compressed(os.path.join(HERE, f"{name}_samples.npz"), idx=idx, counts=sc,
                                canonical=can)
        print(f"{name}: n={xy.shape[0]} oracle {t1 - t0:.1f}s qhull {t2 - t1:.1f}s ok", flush=True)
    sizes = inputs.batch_sizes(256, 1024, 65536, seed=4)
    pts, boxes = inputs.batch_points(sizes, seed=4)
    out["workloads"]["c4"] = {
        "sizes": sizes.tolist(), "total": int(sizes.sum()),
        "points_sha256": inputs.sha256(np.concatenate(pts)),
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
