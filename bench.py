#!/usr/bin/env python3
"""BFS throughput on one MI355X: single-source, batched and served.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-batched] [--no-served]

Prints one JSON line on stdout.  The headline `value` is single-source BFS
GTEPS on R-MAT scale 22 (edge factor 16, a/b/c = .57/.19/.19, seeded
permutation, symmetrised, deduplicated): a step is one BFS from one of 64
seeded sources drawn uniformly from the non-isolated vertices, graph resident
in HBM, levels written to HBM.  GTEPS counts m_reached / 2 undirected edges
per BFS (Graph500 convention).  Beside it:

  batched  R-MAT scale 24, 64 seeded sources in one multi-source BFS vs the
           same 64 sources as 64 single-source BFS: queries/s of both, and
           every source's levels compared with its single-source levels.
  served   8 HBM-resident graphs (R-MAT 16/18/20/22/24, road-like grids of
           1M/4M/16M vertices), Poisson arrivals of (graph, source) at swept
           rates, continuous batching on 4 lanes under a device-memory budget
           vs a plain one-query-at-a-time server: queries/s, p50/p99 latency.
  cpu_baseline  the reference CPU BFS on one pinned host core.

Roofline (algorithmic bytes of one BFS from source s, int32 levels):
    B(s) = 4 n + 28 n_R + 8 m_R
  4 n level init; per reached vertex 16 B of offsets + 12 B of level write,
  enqueue and dequeue; per scanned CSR entry 4 B column + 4 B status probe.
`roofline.achieved` divides the mean B(s) by the mean device time of a BFS
(HIP events around its launch sequence on the handle's stream); the
direction-optimising search reads fewer bytes than B(s) counts, so `frac`
is the fraction of the top-down roofline, as `roofline_bfs_fraction` is.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
SEED = 20260003          # single-source R-MAT s22
SEED_BATCH = 20260004    # batched R-MAT s24 (also the s24 graph of the served mix)
SEED_SERVE = 20260005    # the other served graphs and the query stream
RMAT_ABC = (0.57, 0.19, 0.19)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pick_sources(offsets, k, seed):
    deg = np.diff(offsets)
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def sync(gg):
    gg._ffi.check(gg.lib().gb_device_synchronize(0))


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


def single_source(gg, args):
    t0 = time.time()
    G = gg.Graph.rmat(args.scale, args.edge_factor, *RMAT_ABC, seed=SEED, permute=True)
    n, m = G.n, G.m
    log(f"generated rmat s{args.scale} ef{args.edge_factor}: n={n} m={m} in {time.time() - t0:.1f}s")
    sources = pick_sources(G.offsets, 64, SEED)
    dg = gg.DeviceGraph(G, 0)
    cfg = dg.get_config()
    if args.block:
        cfg["block"] = args.block
    if args.k:
        cfg["edges_per_thread"] = args.k
    if args.mode is not None:
        cfg["mode"] = args.mode
    if args.blocks_per_cu is not None:
        cfg["blocks_per_cu"] = args.blocks_per_cu
    if args.dense_edges:
        cfg["dense_edges"] = args.dense_edges
    dg.config(**cfg)
    if args.alpha or args.beta or args.no_bottom_up:
        dg.direction_config(not args.no_bottom_up, args.alpha or 0, args.beta or 0)
    tuning = None
    if not (args.no_tune or args.block or args.k or args.alpha or args.beta or args.no_bottom_up or args.dense_edges):
        t0 = time.time()
        tuning = dg.autotune(pick_sources(G.offsets, 8, SEED + 1), reps=2)
        log(f"autotune ({time.time() - t0:.1f}s): {tuning}")
    d_levels = gg.DeviceArray(0, np.int32, n)
    log(f"uploaded; config {dg.get_config()}")

    for i in range(args.warmup):
        dg.bfs_device(int(sources[i % len(sources)]), d_levels)
    sync(gg)
    stats = []
    t_begin = time.perf_counter()
    for i in range(args.steps):
        stats.append(dg.bfs_device(int(sources[i % len(sources)]), d_levels))
    sync(gg)
    wall = time.perf_counter() - t_begin
    log(f"timed {args.steps} BFS in {wall * 1e3:.1f} ms")

    edges = sum(s["m_reached"] / 2 for s in stats)
    dev_s = sum(s["ms"] for s in stats) / 1e3
    b_alg = [4 * n + 28 * s["n_reached"] + 8 * s["m_reached"] for s in stats]
    achieved = (sum(b_alg) / len(stats)) / (dev_s / len(stats)) / 1e9
    t_roof = sum(b / (HBM_PEAK_GBS * 1e9) for b in b_alg)
    bu_levels = dg.last_bottom_up()
    per_gteps = [s["m_reached"] / 2 / (s["ms"] / 1e3) / 1e9 for s in stats]

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"bfs_rmat{args.scale}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_bfs")

    result = {
        "metric": "GTEPS single-source BFS + batched queries/s & p99 latency vs plain, same GPU",
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
        "data": "synthetic (R-MAT and grid graphs generated in-process, seeded)",
        "config": {
            "workload": f"rmat-s{args.scale}-ef{args.edge_factor} single-source BFS, 64 seeded sources",
            "scale": args.scale, "edge_factor": args.edge_factor, "abc": list(RMAT_ABC),
            "seed": SEED, "n": n, "m_directed": m, "sources": len(sources), "kernel": dg.get_config(),
            "direction": dg.get_direction_config(),
            "tuning": tuning,
        },
        "device_gteps": round(edges / dev_s / 1e9, 4),
        "gteps_per_source": {"harmonic_mean": round(len(per_gteps) / sum(1 / g for g in per_gteps), 4),
                             "min": round(min(per_gteps), 4), "median": round(float(np.median(per_gteps)), 4)},
        "roofline_bfs_fraction": round(t_roof / dev_s, 4),
        "last_bottom_up_levels": list(bu_levels),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "one BFS: the replayed launch sequence (bfs_clear_kernel, which seeds the source, then "
                      "one bfs_expand_kernel per level, top-down or bottom-up; bfs_compact_kernel only on "
                      "directed graphs), HIP events on its stream",
            "kernel_ms_avg": round(dev_s * 1e3 / len(stats), 4),
            "alg_bytes_per_launch": int(sum(b_alg) / len(b_alg)),
        },
        "cpu_baseline": None,
    }
    return result, G, dg, sources, d_levels


def batched(gg, args):
    """R-MAT s24: one 64-source batch vs 64 single-source BFS, same GPU.  Both
    paths run the configuration the autotuner picks for this graph (tuned on
    other seeded sources than the timed ones)."""
    t0 = time.time()
    G = gg.Graph.rmat(24, 16, *RMAT_ABC, seed=SEED_BATCH, permute=True)
    n = G.n
    log(f"batched: rmat s24 n={n} m={G.m} in {time.time() - t0:.1f}s")
    sources = pick_sources(G.offsets, 64, SEED_BATCH)
    dg = gg.DeviceGraph(G, 0)
    single_tuning = None
    if not args.no_tune:
        t0 = time.time()
        single_tuning = dg.autotune(pick_sources(G.offsets, 8, SEED_BATCH + 2), reps=2)
        log(f"single-source autotune on s24 ({time.time() - t0:.1f}s): {single_tuning}")
    d_rows = gg.DeviceArray(0, np.int32, 64 * n)
    d_lv = gg.DeviceArray(0, np.int32, n)
    reps = args.batch_reps

    def time_batch():
        st = dg.bfs_batch_device(sources, d_rows, stats=True)  # warm-up: workspace + graph capture
        sync(gg)
        t = time.perf_counter()
        for _ in range(reps):
            dg.bfs_batch_device(sources, d_rows)
        sync(gg)
        return st, (time.perf_counter() - t) / reps

    st, batch_s = time_batch()
    default_s, tuning = batch_s, None
    if not args.no_tune:
        t0 = time.time()
        tuning = dg.batch_autotune(pick_sources(G.offsets, 64, SEED_BATCH + 1), reps=2)
        log(f"batch autotune ({time.time() - t0:.1f}s): {tuning}")
        st, batch_s = time_batch()
    for s in sources[:2]:
        dg.bfs_device(int(s), d_lv)
    sync(gg)
    t = time.perf_counter()
    for s in sources:
        dg.bfs_device(int(s), d_lv)
    sync(gg)
    plain_s = time.perf_counter() - t
    mismatches = 0
    for i, s in enumerate(sources):
        dg.bfs_device(int(s), d_lv)
        if not np.array_equal(d_rows.numpy_slice(i * n, n), d_lv.numpy()):
            mismatches += 1
    edges = sum(st["m_reached"][:len(sources)]) / 2
    out = {
        "workload": "rmat-s24-ef16, 64 seeded sources in one multi-source BFS",
        "seed": SEED_BATCH, "n": n, "m_directed": G.m, "sources": len(sources),
        "kernel": dg.get_batch_config(),
        "direction": dg.get_batch_direction_config(),
        "tuning": tuning,
        "single_source_tuning": single_tuning,
        "batch_ms_default_config": round(default_s * 1e3, 3),
        "batch_ms": round(batch_s * 1e3, 3),
        "batched_qps": round(len(sources) / batch_s, 1),
        "batched_gteps": round(edges / batch_s / 1e9, 3),
        "plain_qps": round(len(sources) / plain_s, 1),
        "plain_gteps": round(edges / plain_s / 1e9, 3),
        "speedup_vs_plain": round(plain_s / batch_s, 2),
        "num_levels": st["num_levels"],
        "rows_equal_single_source": len(sources) - mismatches,
        "parity_mismatches": mismatches,
    }
    log(f"batched: {out}")
    return out, G, dg


def served(gg, args, g24=None, dg24=None):
    """8 resident graphs, Poisson query stream, batched server vs plain loop."""
    specs = [("rmat16", lambda: gg.Graph.rmat(16, 16, *RMAT_ABC, seed=SEED_SERVE)),
             ("rmat18", lambda: gg.Graph.rmat(18, 16, *RMAT_ABC, seed=SEED_SERVE)),
             ("rmat20", lambda: gg.Graph.rmat(20, 16, *RMAT_ABC, seed=SEED_SERVE)),
             ("rmat22", lambda: gg.Graph.rmat(22, 16, *RMAT_ABC, seed=SEED_SERVE)),
             ("rmat24", lambda: g24 if g24 is not None else gg.Graph.rmat(24, 16, *RMAT_ABC, seed=SEED_BATCH)),
             ("grid1M", lambda: gg.Graph.grid(1000, 1000, 0.05, seed=SEED_SERVE)),
             ("grid4M", lambda: gg.Graph.grid(2000, 2000, 0.05, seed=SEED_SERVE)),
             ("grid16M", lambda: gg.Graph.grid(4000, 4000, 0.05, seed=SEED_SERVE))]
    gs, dgs, cands = [], [], []
    for name, fn in specs:
        t = time.time()
        g = fn()
        gs.append(g)
        dgs.append(dg24 if (name == "rmat24" and dg24 is not None) else gg.DeviceGraph(g, 0))
        cands.append(np.nonzero(np.diff(g.offsets))[0].astype(np.int32))
        log(f"served: {name} n={g.n} m={g.m} ({time.time() - t:.1f}s)")

    def run(srv, gids, rate, nq, seed):
        rng = np.random.default_rng(seed)
        gi = rng.integers(len(gs), size=nq)
        src = [int(rng.choice(cands[g])) for g in gi]
        gaps = rng.exponential(1.0 / rate, size=nq) if rate else np.zeros(nq)
        tickets = [None] * nq
        failures = []

        def releaser():
            # Results are consumed and released as they complete, in any order.
            pending, nxt_i, finished = [], 0, 0
            while finished < nq:
                while nxt_i < nq and tickets[nxt_i] is not None:
                    pending.append(tickets[nxt_i])
                    nxt_i += 1
                still = []
                for t in pending:
                    rc, _ = srv.poll(t)
                    if rc == gg.GB_EAGAIN:
                        still.append(t)
                        continue
                    if rc != 0:
                        failures.append(rc)
                    srv.release(t)
                    finished += 1
                if len(still) == len(pending):
                    time.sleep(0.0005)
                pending = still

        th = threading.Thread(target=releaser)
        srv.reset_stats()
        th.start()
        nxt = time.perf_counter()
        for i in range(nq):
            nxt += gaps[i]
            while time.perf_counter() < nxt:
                pass
            tickets[i] = srv.submit(gids[gi[i]], src[i])
        th.join()
        st = srv.stats()
        return {"offered_qps": round(rate, 1) if rate else "saturated", "queries": nq,
                "qps": round(st["qps"], 1), "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                "mean_batch": round(st["mean_batch"], 2), "failed": len(failures),
                "memory_used": st["memory_used"]}

    out = {"graphs": [{"name": s[0], "n": g.n, "m_directed": g.m} for s, g in zip(specs, gs)],
           "seed": SEED_SERVE, "memory_budget": args.serve_budget}
    plain = gg.Server(lanes=1, plain=True)
    pids = [plain.add_graph(d) for d in dgs]
    run(plain, pids, 0, 16, 1)
    p_sat = run(plain, pids, 0, args.serve_plain_queries, 2)
    log(f"served plain saturated: {p_sat}")
    batch = gg.Server(lanes=4, max_batch=64, memory_budget=args.serve_budget)
    bids = [batch.add_graph(d) for d in dgs]
    run(batch, bids, 0, 64, 3)
    b_sat = run(batch, bids, 0, args.serve_batch_queries, 4)
    log(f"served batched saturated: {b_sat}")
    sweep = []
    cap = p_sat["qps"]
    for f in (0.5, 1.0, 2.0, 4.0):
        rate = cap * f
        nq = int(max(rate * args.serve_seconds, 24))
        row = {"rate_x_plain": f, "batched": run(batch, bids, rate, nq, 10 + int(f * 4))}
        if f <= 1.0:
            row["plain"] = run(plain, pids, rate, nq, 10 + int(f * 4))
        log(f"served sweep: {row}")
        sweep.append(row)
    plain.close()
    batch.close()
    out.update({"plain_saturated": p_sat, "batched_saturated": b_sat, "sweep": sweep,
                "saturated_speedup_vs_plain": round(b_sat["qps"] / max(p_sat["qps"], 1e-9), 2)})
    return out


def copy_probe(bytes_):
    """A plain device copy of known size, for scaling the HBM byte counters."""
    import torch
    x = torch.empty(bytes_ // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": x.numel() * 4, "copy_probe_bytes_written": y.numel() * 4}))


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
    ap.add_argument("--blocks-per-cu", type=int, default=None)
    ap.add_argument("--dense-edges", type=int, default=None, help="frontier edges above which a level compacts a bitmap")
    ap.add_argument("--alpha", type=int, default=None, help="bottom-up switch: frontier edges * alpha > unexplored")
    ap.add_argument("--beta", type=int, default=None, help="back top-down below n / beta frontier vertices")
    ap.add_argument("--no-bottom-up", action="store_true", help="top-down only")
    ap.add_argument("--no-tune", action="store_true", help="keep the default kernel configuration")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--serve-budget", type=int, default=64 << 30,
                    help="device bytes for the batched server's workspaces and result buffers")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="arrival window of each swept rate")
    ap.add_argument("--serve-plain-queries", type=int, default=96)
    ap.add_argument("--serve-batch-queries", type=int, default=2048)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (byte-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import gpu_graphs_amd as gg
    assert gg.device_count() > 0, "no GPU visible"

    result, G, dg, sources, d_levels = single_source(gg, args)
    if not args.no_cpu_baseline:
        def gpu_levels(s):
            dg.bfs_device(s, d_levels)
            return d_levels.numpy()
        result["cpu_baseline"] = cpu_baseline(G.offsets, G.cols, sources, gpu_levels, args.cpu_budget)
    del dg, d_levels, G
    g24 = dg24 = None
    if not args.no_batched:
        result["batched"], g24, dg24 = batched(gg, args)
    if not args.no_served:
        result["served"] = served(gg, args, g24, dg24)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
