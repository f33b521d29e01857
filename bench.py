#!/usr/bin/env python3
"""Headline benchmark: points/sec of the global periodic Delaunay
triangulation of 1,048,576 uniform points (box 1024 x 1024, cell size 1) on
one MI355X, FP32-filtered predicates with FP64 fallback and all-FP64 timed
side by side (rows identical).  The headline value is the library's default
path, FP64; the FP32-filtered rate is reported beside it.

A step is one DelaunayGPU.globalDelaunayTriangulation call (counting-sort
cell list + one-ring kernel + the status readback of the C ABI) on points
already resident in HBM.  The timed region covers exactly --steps calls,
bracketed by device synchronisation.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_POINTS = 1 << 20
BOX = 1024.0
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (FP64 coordinates and int32 indices in HBM, C = N cells):
#   whole triangulation 84 N + 12 C; one-ring tile kernel 52 B per point
#   (16 B cell offset + 4 B caller index read, 4 B cell start read, 24 B of
#   ring entries (6 per point on the torus) + 4 B count written).
RING_BYTES_PER_POINT = 52.0
TRI_BYTES = lambda n, c: 84.0 * n + 12.0 * c  # noqa: E731


def md_config(args):
    if not args.md_config:
        return None
    block, ts, rl = (int(v) for v in args.md_config.split(","))
    return {"block": block, "tile_shift": ts, "repair_lanes": rl}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_POINTS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--block", type=int, default=None, help="threads per block (one-ring kernel)")
    ap.add_argument("--tiled", type=int, default=None, help="1: 32x32-cell tiles (LDS kernel), 2: the same with Morton order inside, 0: row-major")
    ap.add_argument("--cell-scale", type=float, default=None, help="cell size multiplier")
    ap.add_argument("--tile-shift", type=int, default=None,
                    help="LDS tile edge 2^k cells (4 or 5; default: automatic)")
    ap.add_argument("--lanes", type=int, default=None,
                    help="lanes per point of the tile kernel (1, 2, 4 or 64; default: automatic)")
    ap.add_argument("--no-small-n", action="store_true",
                    help="skip the 16,384 / 65,536 / 262,144-point timings")
    ap.add_argument("--precisions", default="fp32f,fp64", help="comma list of fp32f,fp64")
    ap.add_argument("--md-steps", type=int, default=1000,
                    help="steps of the 256-set Brownian test-and-repair loop (0: skip)")
    ap.add_argument("--md-precision", default="fp64", choices=("fp64", "fp32f"),
                    help="predicates of the Brownian loop's repairs")
    ap.add_argument("--md-config", default=None, metavar="BLOCK,TILE_SHIFT,REPAIR_LANES",
                    help="fixed configuration of the Brownian loop instead of the tuner's")
    ap.add_argument("--no-voronoi", action="store_true", help="skip the Voronoi energy timing")
    ap.add_argument("--no-general", action="store_true",
                    help="skip the general (sheared) box timing")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 4,096-point case (GPU and CPU oracle timings)")
    ap.add_argument("--no-exact", action="store_true",
                    help="skip the exact precision mode's timings")
    ap.add_argument("--no-c3b", action="store_true",
                    help="skip the jittered-lattice case (2^20 points, both precisions)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the bench line, run a plain device copy of MIB MiB three "
                         "times (reference kernel for scaling rocprofv3 FETCH/WRITE_SIZE)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_one_ring_tile_uniform1M_hbm_traffic.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE passes of this workload, scaled by the copy
    calibration); None when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def time_path(dg, pts, tri, num, steps, warmup, torch):
    for _ in range(warmup):
        dg.globalDelaunayTriangulation(pts, tri, num)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the library sums its own HIP-event kernel times over the timed calls;
    # the loop itself makes no other call
    s0 = dg.stats()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        dg.globalDelaunayTriangulation(pts, tri, num)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / steps
    s1 = dg.stats()
    calls = s1["profiled_calls"] - s0["profiled_calls"]       # steps, or 0 without profile = 1
    assert calls in (0, steps)
    per = {k: (s1[k + "_total"] - s0[k + "_total"]) / max(calls, 1)
           for k in ("ms_ring", "ms_cells", "ms_defer")}
    return ms, wall, per["ms_ring"], per["ms_cells"], per["ms_defer"], s1


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry), or 'unknown'."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_c1(torch, cpu_runs, with_cpu):
    """4,096 uniform points in a box of side 64 (the reference's CPU-runnable
    case): GPU points/s per precision (HIP events, 200 calls after 200, so that
    the clocks have ramped up for calls of a few tens of microseconds) and,
    unless skipped, the CPU oracle on the same points (median of cpu_runs)."""
    from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs
    from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries
    n, L = 4096, 64.0
    xy = inputs.uniform_points(n, L, L, seed=1)
    pts = torch.from_numpy(xy).cuda()
    res = {"workload": "4096 uniform points, box 64, cell 1.0"}
    for name, prec in (("fp32f", PREC_FP32F), ("fp64", PREC_FP64)):
        dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(L, L))
        dg.setPrecision(prec)
        tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
        num = torch.zeros(n, dtype=torch.int32, device="cuda")
        ms = time_path(dg, pts, tri, num, 200, 200, torch)[0]
        res[f"gpu_{name}_points_per_s"] = n / (ms * 1e-3)
    if with_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.triangulate(xy, L, L, 1.0, 16, nthreads=threads)
        times = []
        for _ in range(cpu_runs):
            t0 = time.perf_counter()
            oracle.triangulate(xy, L, L, 1.0, 16, nthreads=threads)
            times.append(time.perf_counter() - t0)
        res["cpu_points_per_s"] = n / sorted(times)[len(times) // 2]
        res["cpu_cores"] = threads
    return res


def bench_c3b(torch, steps, warmup):
    """The jittered triangular lattice: 1024 x 1024 points, spacing 1, rows
    sqrt(3)/2 apart, Gaussian jitter sigma = 1e-5 (near-cocircular quads),
    box 1024 x 886.81, cell size 1: points/s and fallback fraction per
    precision, rows identical across precisions and equal to the committed
    oracle hash."""
    from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings
    from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries
    lx, ly = inputs.lattice_box(1024, 1024)
    xy = inputs.lattice_points(1024, 1024)
    n = xy.shape[0]
    pts = torch.from_numpy(xy).cuda()
    res = {"workload": f"jittered triangular lattice 1024x1024 (sigma 1e-5), box {lx:g}x{ly:.6f}, "
                       f"cell 1.0"}
    hashes = {}
    for name, prec in (("fp32f", PREC_FP32F), ("fp64", PREC_FP64)):
        dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(lx, ly))
        dg.setPrecision(prec)
        dg.ctx.set_config(profile=1)
        tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
        num = torch.zeros(n, dtype=torch.int32, device="cuda")
        ms, _, ring_ms, _, _, st = time_path(dg, pts, tri, num, steps, warmup, torch)
        hashes[name] = rings.ring_hash(tri.cpu().numpy().reshape(n, dg.MaxSize), num.cpu().numpy())
        res[f"{name}_points_per_s"] = n / (ms * 1e-3)
        res[f"{name}_ms_per_step"] = ms
        res[f"{name}_one_ring_ms"] = ring_ms
        res[f"{name}_deferred_points"] = st["deferred"]
        if name == "fp32f":
            res["fallback_fraction"] = st["fallbacks"] / max(1, st["predicates"])
            res["predicates_per_point"] = st["predicates"] / n
    res["topology_identical_fp32_fp64"] = hashes["fp32f"] == hashes["fp64"]
    gpath = os.path.join(ROOT, "tests", "golden", "golden.json")
    with open(gpath) as fh:
        res["matches_golden_oracle_hash"] = (json.load(fh)["workloads"]["c3b"]["ring_sha256"]
                                             == hashes["fp32f"])
    return res


def bench_exact(torch, steps, warmup):
    """The exact precision mode (every decision exact): points/s on the 2^20
    uniform workload (rings must equal the committed oracle hash) and on a
    near-degenerate lattice, 1024 x 1024 centres of a square lattice of
    spacing 0.1 (rounded to FP64: every square's four points nearly
    cocircular), box 102.4, cell 0.1, with the FP64 path beside it (its rings
    there are not a triangulation: their counts need not sum to 6N)."""
    import numpy as np
    from gpudelaunaytriangulation_amd import PREC_EXACT, PREC_FP64, _lib, inputs, rings
    from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries
    res = {}
    n = N_POINTS
    xy = inputs.uniform_points(n, BOX, BOX, seed=2)
    dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(BOX, BOX))
    dg.setPrecision(PREC_EXACT)
    pts = torch.from_numpy(xy).cuda()
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    ms, *_ = time_path(dg, pts, tri, num, steps, warmup, torch)
    h = rings.ring_hash(tri.cpu().numpy().reshape(n, dg.MaxSize), num.cpu().numpy())
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        golden = json.load(fh)["workloads"]["c2"]["ring_sha256"]
    res["uniform_1M_points_per_s"] = n / (ms * 1e-3)
    res["uniform_1M_ms_per_step"] = ms
    res["uniform_1M_matches_golden_oracle_hash"] = h == golden
    m, hsp = 1024, 0.1
    X, Y = np.meshgrid(np.arange(m), np.arange(m), indexing="xy")
    lat = np.ascontiguousarray(np.stack([(X.reshape(-1) + 0.5) * hsp, (Y.reshape(-1) + 0.5) * hsp], 1))
    L = m * hsp
    pts = torch.from_numpy(lat).cuda()
    res["lattice_workload"] = f"1024x1024 square lattice of spacing 0.1 (FP64), box {L:g}, cell 0.1"
    for name, prec in (("exact", PREC_EXACT), ("fp64", PREC_FP64)):
        dg = DelaunayGPU(lat.shape[0], 16, hsp, PeriodicBoundaries(L, L))
        dg.setPrecision(prec)
        tri = torch.zeros(lat.shape[0] * dg.MaxSize, dtype=torch.int32, device="cuda")
        num = torch.zeros(lat.shape[0], dtype=torch.int32, device="cuda")
        try:
            ms, *_ = time_path(dg, pts, tri, num, steps, warmup, torch)
            res[f"lattice_{name}_points_per_s"] = lat.shape[0] / (ms * 1e-3)
            res[f"lattice_{name}_sum_counts_eq_6N"] = int(num.sum().item()) == 6 * lat.shape[0]
        except _lib.GdtError as e:
            res[f"lattice_{name}_error"] = str(e)
    return res


def bench_general(torch, steps, warmup):
    """A general (triclinic) box: 2^20 uniform points in the sheared box with
    lattice vectors (1024, 0) and (320, 1024) (density 1, cell size 1), FP64
    (general boxes run the row-major per-point kernels): points/s of the
    global triangulation, time of a test-and-repair after a Gaussian move of
    sigma 0.01 (wrapped by whole lattice vectors), and whether the rings equal
    the CPU mode's (the same routines on the host)."""
    import ctypes
    import numpy as np
    from gpudelaunaytriangulation_amd import PREC_FP64, _lib, inputs, rings
    from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries
    box = (1024.0, 320.0, 0.0, 1024.0)
    n = 1 << 20
    xy = inputs.general_points(n, box, seed=5)
    pts = torch.from_numpy(xy).cuda()
    dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(*box))
    dg.setPrecision(PREC_FP64)
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    ms, _, _, _, _, st = time_path(dg, pts, tri, num, steps, warmup, torch)
    nb = tri.cpu().numpy().reshape(n, dg.MaxSize)
    cnt = num.cpu().numpy()
    L = _lib.load()
    hn = np.full((n, dg.MaxSize), -1, dtype=np.int32)
    hc = np.zeros(n, dtype=np.int32)
    req = ctypes.c_int(0)
    bp = (ctypes.c_double * 4)(*box)
    hst = L.gdt_host_triangulate_general(xy.ctypes.data, n, bp, 1.0, dg.MaxSize, hn.ctypes.data,
                                         hc.ctypes.data, PREC_FP64, 0, ctypes.byref(req))
    moved = inputs.wrap_into_general_box(xy + 0.01 * inputs.gaussian(6, 2 * n).reshape(n, 2), box)
    mpts = torch.from_numpy(moved).cuda()
    times = []
    for _ in range(5):
        t = tri.clone()
        c = num.clone()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dg.testAndRepairDelaunayTriangulation(mpts, t, c)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return {
        "workload": "2^20 uniform points, general box a=(1024,0) b=(320,1024), cell 1.0, fp64",
        "points_per_s": n / (ms * 1e-3),
        "ms_per_step": ms,
        "deferred_points": st["deferred"],
        "cells": [st["xsize"], st["ysize"]],
        "test_and_repair_ms": float(np.median(times)),
        "repaired_points": dg.last_repaired,
        "euler_ok": int(cnt.sum()) == 6 * n,
        "rings_equal_cpu_mode": hst == _lib.GDT_OK and rings.ring_hash(nb, cnt) == rings.ring_hash(hn, hc),
    }


def bench_small_n(torch, sizes=(16384, 65536, 262144)):
    """Uniform points at density 1 (box sqrt(N)), default (automatic)
    configuration: points/s per precision over 200 calls after 200 (HIP events),
    with the tile edge and lanes per point the library chose."""
    from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs
    from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries
    import numpy as np
    out = {}
    for n in sizes:
        L = float(np.sqrt(n))
        pts = torch.from_numpy(inputs.uniform_points(n, L, L, seed=2)).cuda()
        row = {}
        for name, prec in (("fp32f", PREC_FP32F), ("fp64", PREC_FP64)):
            dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(L, L))
            dg.setPrecision(prec)
            tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
            num = torch.zeros(n, dtype=torch.int32, device="cuda")
            ms, _, _, _, _, st = time_path(dg, pts, tri, num, 200, 200, torch)
            row[f"{name}_points_per_s"] = n / (ms * 1e-3)
            row["tile_cells"] = 1 << st["tile_shift"]
            row["lanes_per_point"] = st["lanes"]
        out[str(n)] = row
    return out


def bench_md(steps, torch, precision="fp64", config=None):
    """256 sets, N log-uniform in [1024, 65536], box sqrt(N): one batched
    triangulation, then `steps` hipGraph-replayed steps of sigma = 0.01
    displacement -> test -> repair."""
    import numpy as np
    from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings
    from gpudelaunaytriangulation_amd.delaunay import BatchedDelaunayGPU
    sizes = inputs.batch_sizes(256)
    pts, boxes = inputs.batch_points(sizes)
    bd = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], 16, 1.0)
    bd.setPrecision(PREC_FP32F if precision == "fp32f" else PREC_FP64)
    X = torch.from_numpy(np.concatenate(pts)).cuda()
    tri = torch.zeros(bd.N * bd.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    bd.globalDelaunayTriangulation(X, tri, num)          # also sizes MaxSize
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bd.globalDelaunayTriangulation(X, tri, num)
    torch.cuda.synchronize()
    t_tri = time.perf_counter() - t0
    max_ring0 = int(num.max().item())
    # the loop's configuration (tile edge, lanes per repaired point, block),
    # tuned on 20 replayed steps of the same start (autotune.tune_md)
    # (--md-config fixes it instead, e.g. for a profile of the steady state)
    from gpudelaunaytriangulation_amd import autotune
    if config is None:
        tuned = autotune.tune_md(bd, X, tri, num, steps=20, seed=7)
    else:
        tuned = dict(config, table=[])
        bd.ctx.set_config(**config)
    bd.md_run(X, tri, num, 5, 0.01, seed=99, step0=0, use_graph=True, check_every=5)   # warm-up
    torch.cuda.synchronize()      # (the warm-up also widens the rows to the longest ring + 4)
    t0 = time.perf_counter()
    rep = bd.md_run(X, tri, num, steps, 0.01, seed=99, step0=5, use_graph=True, check_every=100)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    # the final rings against a fresh global triangulation of the final positions
    fresh = BatchedDelaunayGPU(sizes, [(L, L) for L in boxes], bd.MaxSize, 1.0)
    tri2 = torch.zeros(bd.N * fresh.MaxSize, dtype=torch.int32, device="cuda")
    num2 = torch.zeros(bd.N, dtype=torch.int32, device="cuda")
    fresh.globalDelaunayTriangulation(X, tri2, num2)
    torch.cuda.synchronize()
    final_match = (rings.ring_hash(tri.cpu().numpy().reshape(bd.N, bd.MaxSize), num.cpu().numpy())
                   == rings.ring_hash(tri2.cpu().numpy().reshape(bd.N, fresh.MaxSize),
                                      num2.cpu().numpy()))
    return {"workload": f"256 sets, {bd.N} points, sigma 0.01, hipGraph replay, {precision}",
            "points": bd.N, "batch_triangulation_sets_per_s": 256 / t_tri,
            "batch_triangulation_points_per_s": bd.N / t_tri, "steps": rep["steps"],
            "steps_per_s": rep["steps"] / t, "sets_steps_per_s": 256 * rep["steps"] / t,
            "repaired_fraction": rep["flagged_total"] / (rep["steps"] * bd.N),
            "max_ring_start": max_ring0, "max_size": bd.MaxSize,
            "recoveries": rep["recoveries"], "recovered": rep["recovered"],
            "skipped_replays_after_halt": rep["skipped_replays"],
            "final_rings_match": bool(final_match),
            "tuned_config": {k: tuned[k] for k in ("block", "tile_shift", "repair_lanes")},
            "tuner_table_ms_per_step": [(r["block"], r["tile_shift"], r["repair_lanes"],
                                         round(r["ms_per_step"], 4)) for r in tuned["table"]],
            "step_roofline_frac": (48 + 32 + 96 * rep["flagged_total"] / (rep["steps"] * bd.N))
            * bd.N * rep["steps"] / t / 1e9 / HBM_PEAK_GBS}


def bench_voronoi(torch):
    """Energy + hand-written adjoint, 65,536 points (box 256), FP64 and FP32."""
    import numpy as np
    from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs
    from gpudelaunaytriangulation_amd.delaunay import Context, DelaunayGPU, PeriodicBoundaries
    from gpudelaunaytriangulation_amd.voronoi import voronoi_adjoint, voronoi_forward
    L, n = 256.0, 65536
    xy = torch.from_numpy(inputs.uniform_points(n, L, L, seed=5)).cuda()
    dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(L, L))
    tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
    num = torch.zeros(n, dtype=torch.int32, device="cuda")
    dg.globalDelaunayTriangulation(xy, tri, num)
    ctx = Context(0)
    out = {}
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32F)):
        def once():
            a, p, e = voronoi_forward(ctx, xy, tri, num, (L, L), precision=prec)
            return voronoi_adjoint(ctx, xy, tri, num, (L, L), a, p, precision=prec)
        once()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            once()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        out[name] = {"ms_forward_adjoint": ms, "points_per_s": n / ms * 1e3,
                     "roofline_frac": 60.0 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out[name]["_grad"] = once().double()
    # measured gradient errors (topology fixed): the FP64 adjoint against
    # central differences of the FP64 energy (h = 1e-6) on 64 sampled
    # coordinates, and the FP32 adjoint against the FP64 adjoint; both relative
    # to the largest gradient entry
    g64, g32 = out["fp64"].pop("_grad"), out["fp32"].pop("_grad")
    gmax = g64.abs().max().item()
    h = 1e-6
    idx = np.random.default_rng(0).choice(2 * n, size=64, replace=False)
    fd_err = 0.0
    for k in idx:
        i, d = divmod(int(k), 2)
        e = []
        for sgn in (1.0, -1.0):
            xp = xy.clone()
            xp[i, d] += sgn * h
            e.append(voronoi_forward(ctx, xp, tri, num, (L, L), precision=PREC_FP64)[2].item())
        fd = (e[0] - e[1]) / (2 * h)
        fd_err = max(fd_err, abs(fd - g64[i, d].item()))
    out["grad_rel_err_fp64_vs_central_differences"] = fd_err / gmax
    out["grad_rel_err_fp32_vs_fp64"] = (g32 - g64).abs().max().item() / gmax
    out["workload"] = "65536 uniform points, box 256, E = sum (A-1)^2 + (P-3.8)^2"
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from gpudelaunaytriangulation_amd import PREC_FP32F, PREC_FP64, inputs, rings
    from gpudelaunaytriangulation_amd.delaunay import DelaunayGPU, PeriodicBoundaries

    n = args.n
    L = BOX if n == N_POINTS else float(np.sqrt(n))
    xy = inputs.uniform_points(n, L, L, seed=2)
    pts = torch.from_numpy(xy).cuda()
    out = {}
    wanted = args.precisions.split(",")
    for name, prec in (("fp32f", PREC_FP32F), ("fp64", PREC_FP64)):
        if name not in wanted and not (name == "fp32f"):
            continue
        dg = DelaunayGPU(n, 16, 1.0, PeriodicBoundaries(L, L))
        dg.setPrecision(prec)
        dg.ctx.set_config(profile=1, block=args.block, tiled=args.tiled, cell_scale=args.cell_scale,
                          tile_shift=args.tile_shift, lanes=args.lanes)
        tri = torch.zeros(n * dg.MaxSize, dtype=torch.int32, device="cuda")
        num = torch.zeros(n, dtype=torch.int32, device="cuda")
        ms, wall, ring_ms, cell_ms, defer_ms, st = time_path(dg, pts, tri, num, args.steps,
                                                             args.warmup, torch)
        nb = tri.cpu().numpy().reshape(n, dg.MaxSize)
        cnt = num.cpu().numpy()
        out[name] = dict(ms=ms, wall=wall, ring_ms=ring_ms, cell_ms=cell_ms, defer_ms=defer_ms,
                         st=st, hash=rings.ring_hash(nb, cnt), sum=int(cnt.sum()), nb=nb, cnt=cnt)
        print(f"[bench] {name}: {ms:.3f} ms/step (one-ring {ring_ms:.3f} ms, deferred "
              f"{defer_ms:.3f} ms, cell list {cell_ms:.3f} ms), {n / ms * 1e3 / 1e9:.3f} Gpoints/s, "
              f"deferred {st['deferred']}, pred/pt {st['predicates'] / n:.1f}", file=sys.stderr,
              flush=True)

    f32 = out["fp32f"]
    f64 = out.get("fp64", f32)
    identical = (f32["hash"] == f64["hash"] and np.array_equal(f32["nb"], f64["nb"])
                 and np.array_equal(f32["cnt"], f64["cnt"]))
    golden_ok = None
    gpath = os.path.join(ROOT, "tests", "golden", "golden.json")
    if n == N_POINTS and os.path.exists(gpath):
        with open(gpath) as fh:
            golden_ok = json.load(fh)["workloads"]["c2"]["ring_sha256"] == f32["hash"]

    # The headline is the library's default path, FP64 (when measured); the
    # FP32-filtered path, decision-identical, is reported beside it.
    head = f64 if "fp64" in out else f32
    head_name = "fp64" if "fp64" in out else "fp32f"
    head_kernel = "k_one_ring_tile<64, 1>" if head_name == "fp64" else "k_one_ring_tile<32, 1>"
    value = n / (head["ms"] * 1e-3)
    ring_bytes = RING_BYTES_PER_POINT * n
    achieved = ring_bytes / (head["ring_ms"] * 1e-3) / 1e9
    st = head["st"]

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.triangulate(xy, L, L, 1.0, 16, nthreads=threads)     # warm-up
        times = []
        for _ in range(args.cpu_runs):
            t0 = time.perf_counter()
            oracle.triangulate(xy, L, L, 1.0, 16, nthreads=threads)
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        cpu = {"value": n / t, "unit": "points/s", "cores": threads, "kind": "port",
               "sample": f"the same {n} points, full triangulation incl. cell list, "
                         f"median of {args.cpu_runs} runs after 1 warm-up, "
                         f"{threads} OpenMP threads (oracle/gdt_oracle.c) on {cpu_model()}"}

    line = {
        "metric": "points/sec triangulated, 1M uniform: FP32-filtered vs FP64 GPU, topology exact",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_name == "fp64" else "fp32+fp64 fallback (coordinates fp64)",
        "data": "synthetic: SplitMix64 uniform points, seed 2",
        "config": {"workload": f"uniform periodic {n} points, box {L:g}x{L:g}, cell 1.0, "
                               + ("FP64 global triangulation (the library default; the "
                                  "FP32-filtered path is reported beside it)"
                                  if head_name == "fp64" else "FP32-filtered global triangulation"),
                   "headline_path": head_name,
                   "n": n, "box": L, "cell_size": 1.0, "block": dg.ctx.config().block,
                   "tile_cells": 1 << st["tile_shift"], "lanes_per_point": st["lanes"]},
        "fp64_points_per_s": n / (f64["ms"] * 1e-3),
        "fp64_ms_per_step": f64["ms"],
        "fp32f_points_per_s": n / (f32["ms"] * 1e-3),
        "fp32f_ms_per_step": f32["ms"],
        "fp32_over_fp64": f64["ms"] / f32["ms"],
        "topology_identical_fp32_fp64": bool(identical),
        "matches_golden_oracle_hash": golden_ok,
        "sum_counts_eq_6N": f32["sum"] == 6 * n,
        "fallback_fraction": st["fallbacks"] / max(1, st["predicates"]),
        "predicates_per_point": st["predicates"] / n,
        "deferred_points": st["deferred"],
        "kernel_ms": {"one_ring_tile_fp32f": f32["ring_ms"], "one_ring_tile_fp64": f64["ring_ms"],
                      "deferred_fp32f": f32["defer_ms"], "deferred_fp64": f64["defer_ms"],
                      "cell_list": f32["cell_ms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("k_one_ring_tile<64>" if head_name == "fp64"
                                                  else "k_one_ring_tile<32>")
                                 if n == N_POINTS else None),
                     "traffic_source": "profiles/k_one_ring_tile_uniform1M_hbm_traffic.md",
                     "kernel": head_kernel, "algorithmic_bytes_per_launch": ring_bytes},
        "triangulation_roofline_frac": TRI_BYTES(n, n) / (head["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
    }
    if not args.no_c1:
        try:
            line["c1_4096"] = bench_c1(torch, args.cpu_runs, not args.no_cpu_baseline)
        except Exception as e:  # reported, never hides the headline number
            line["c1_4096"] = {"error": repr(e)}
        print(f"[bench] c1: {line['c1_4096']}", file=sys.stderr, flush=True)
    if not args.no_small_n:
        try:
            line["small_n"] = bench_small_n(torch)
        except Exception as e:  # reported, never hides the headline number
            line["small_n"] = {"error": repr(e)}
        print(f"[bench] small_n: {line['small_n']}", file=sys.stderr, flush=True)
    if not args.no_exact:
        try:
            line["exact_mode"] = bench_exact(torch, args.steps, args.warmup)
        except Exception as e:  # reported, never hides the headline number
            line["exact_mode"] = {"error": repr(e)}
        print(f"[bench] exact mode: {line['exact_mode']}", file=sys.stderr, flush=True)
    if not args.no_c3b:
        try:
            line["c3b"] = bench_c3b(torch, args.steps, args.warmup)
        except Exception as e:  # reported, never hides the headline number
            line["c3b"] = {"error": repr(e)}
        print(f"[bench] c3b: {line['c3b']}", file=sys.stderr, flush=True)
    if not args.no_general:
        try:
            line["general_box"] = bench_general(torch, args.steps, args.warmup)
        except Exception as e:  # reported, never hides the headline number
            line["general_box"] = {"error": repr(e)}
        print(f"[bench] general box: {line['general_box']}", file=sys.stderr, flush=True)
    if args.md_steps > 0:
        try:
            line["md_c4"] = bench_md(args.md_steps, torch, args.md_precision, md_config(args))
        except Exception as e:  # reported, never hides the headline number
            line["md_c4"] = {"error": repr(e)}
        print(f"[bench] md: {line['md_c4']}", file=sys.stderr, flush=True)
    if not args.no_voronoi:
        try:
            line["voronoi_c5"] = bench_voronoi(torch)
        except Exception as e:
            line["voronoi_c5"] = {"error": repr(e)}
        print(f"[bench] voronoi: {line['voronoi_c5']}", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)

    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 3 x {nbytes} bytes read + written", file=sys.stderr,
              flush=True)


if __name__ == "__main__":
    main()
