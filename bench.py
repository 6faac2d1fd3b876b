#!/usr/bin/env python3
"""Single-source BFS throughput on one MI355X (R-MAT scale 22, edge factor 16).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one BFS from one of 64 seeded sources (drawn uniformly from the
non-isolated vertices), graph resident in HBM, levels written to HBM.  Prints
one JSON line on stdout: GTEPS over the timed steps (Graph500 counting:
m_reached / 2 undirected edges per BFS), the HBM roofline of the traversal
kernel, and the CPU BFS baseline timed on this host.

Roofline (algorithmic bytes of one BFS from source s, int32 levels):
    B(s) = 4 n + 28 n_R + 8 m_R
  4 n level init; per reached vertex 16 B of offsets + 12 B of level write,
  enqueue and dequeue; per scanned CSR entry 4 B column + 4 B status probe.
The traversal kernel (everything but the 4 n init) moves 28 n_R + 8 m_R.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
SEED = 20260003


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pick_sources(offsets, k, seed):
    deg = np.diff(offsets)
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def cpu_baseline(offsets, cols, sources, gpu_levels_fn, budget_s):
    """Times the reference CPU BFS (or the oracle port) on one core."""
    from oracle import oracle
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    use_ref = oracle.ref_available() and offsets[-1] < 2**31
    total_edges, total_s, done, mismatches = 0, 0.0, 0, 0
    t_start = time.time()
    for s in sources:
        if use_ref:
            lv, _, wall_ms, _ = oracle.ref_cpu_bfs(offsets, cols, int(s))
            ms = wall_ms
            m_r = int(np.diff(offsets)[lv >= 0].sum())
        else:
            lv, ms, _, m_r = oracle.bfs_timed(offsets, cols, int(s))
        total_edges += m_r / 2
        total_s += ms / 1e3
        done += 1
        if not np.array_equal(lv, gpu_levels_fn(int(s))):
            mismatches += 1
        log(f"cpu baseline: source {int(s)} {ms:.1f} ms")
        if time.time() - t_start > budget_s:
            break
    try:
        os.sched_setaffinity(0, set(range(os.cpu_count() or 1)))
    except (AttributeError, OSError):
        pass
    return {
        "value": round(total_edges / total_s / 1e9, 6),
        "unit": "GTEPS",
        "cores": 1,
        "kind": "reference" if use_ref else "port",
        "sample": f"{done} of the {len(sources)} sources, serial BFS on one pinned core "
                  f"({'reference cpuBFS::cpuBFS built from its sources' if use_ref else 'oracle restatement'}); "
                  f"host has {os.cpu_count()} cpus",
        "parity_mismatches": mismatches,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--k", type=int, default=None)
    ap.add_argument("--mode", type=int, default=None)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import gpu_graphs_amd as gg
    assert gg.device_count() > 0, "no GPU visible"

    t0 = time.time()
    G = gg.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=SEED, permute=True)
    n, m = G.n, G.m
    offsets, cols = G.offsets, G.cols
    log(f"generated rmat s{args.scale} ef{args.edge_factor}: n={n} m={m} in {time.time() - t0:.1f}s")
    sources = pick_sources(offsets, 64, SEED)
    dg = gg.DeviceGraph(G, 0)
    cfg = dg.get_config()
    if args.block:
        cfg["block"] = args.block
    if args.k:
        cfg["edges_per_thread"] = args.k
    if args.mode is not None:
        cfg["mode"] = args.mode
    dg.config(**cfg)
    d_levels = gg.DeviceArray(0, np.int32, n)
    log(f"uploaded; config {dg.get_config()}")

    for i in range(args.warmup):
        dg.bfs_device(int(sources[i % len(sources)]), d_levels)
    gg._ffi.check(gg.lib().gb_device_synchronize(0))

    stats = []
    t_begin = time.perf_counter()
    for i in range(args.steps):
        stats.append(dg.bfs_device(int(sources[i % len(sources)]), d_levels))
    gg._ffi.check(gg.lib().gb_device_synchronize(0))
    wall = time.perf_counter() - t_begin
    log(f"timed {args.steps} BFS in {wall * 1e3:.1f} ms")

    edges = sum(s["m_reached"] / 2 for s in stats)
    dev_s = sum(s["ms"] for s in stats) / 1e3
    ker_s = sum(s["kernel_ms"] for s in stats) / 1e3
    b_alg = [4 * n + 28 * s["n_reached"] + 8 * s["m_reached"] for s in stats]
    b_ker = [28 * s["n_reached"] + 8 * s["m_reached"] for s in stats]
    achieved = (sum(b_ker) / len(stats)) / (ker_s / len(stats)) / 1e9
    t_roof = sum(b / (HBM_PEAK_GBS * 1e9) for b in b_alg)
    per_gteps = [s["m_reached"] / 2 / (s["ms"] / 1e3) / 1e9 for s in stats]

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"bfs_persistent_rmat{args.scale}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    result = {
        "metric": "GTEPS single-source BFS (R-MAT scale 22, 64 seeded sources)",
        "value": round(edges / wall / 1e9, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (R-MAT generated in-process, seeded)",
        "config": {
            "workload": f"rmat-s{args.scale}-ef{args.edge_factor} single-source top-down BFS",
            "scale": args.scale, "edge_factor": args.edge_factor, "abc": [0.57, 0.19, 0.19],
            "seed": SEED, "n": n, "m_directed": m, "sources": len(sources), "kernel": dg.get_config(),
        },
        "device_gteps": round(edges / dev_s / 1e9, 4),
        "gteps_per_source": {"harmonic_mean": round(len(per_gteps) / sum(1 / g for g in per_gteps), 4),
                             "min": round(min(per_gteps), 4), "median": round(float(np.median(per_gteps)), 4)},
        "roofline_bfs_fraction": round(t_roof / dev_s, 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "bfs_persistent_kernel",
            "kernel_ms_avg": round(ker_s * 1e3 / len(stats), 4),
            "alg_bytes_per_launch": int(sum(b_ker) / len(b_ker)),
        },
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        def gpu_levels(s):
            dg.bfs_device(s, d_levels)
            return d_levels.numpy()
        result["cpu_baseline"] = cpu_baseline(offsets, cols, sources, gpu_levels, args.cpu_budget)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
