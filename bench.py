#!/usr/bin/env python3
"""BFS throughput on one MI355X: single-source, road grid, batched and served.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-grid] [--no-batched] [--no-served]
                    [--retune] [--tuning-out PATH]

Prints one JSON line on stdout.  The headline `value` is single-source BFS
GTEPS on R-MAT scale 22 (edge factor 16, a/b/c = .57/.19/.19, seeded
permutation, symmetrised, deduplicated): a step is one pass over the 64
seeded sources (64 BFS, one after the other), graph resident in HBM, levels
written to HBM.  GTEPS counts m_reached / 2 undirected edges per BFS
(Graph500 convention).  Beside it:

  grid     the 7000 x 7000 road-like grid (49M vertices, spanning tree + 5%
           of the other grid edges), 16 seeded sources: GTEPS, levels and
           microseconds per level, the reference CPU BFS on the same sources.
  batched  R-MAT scale 24, 64 seeded sources in one multi-source BFS vs the
           same 64 sources as 64 single-source BFS: queries/s of both, and
           every source's levels compared with its single-source levels.
  served   8 HBM-resident graphs (R-MAT 16/18/20/22/24, road-like grids of
           1M/4M/16M vertices), Poisson arrivals of (graph, source) at swept
           rates, continuous batching on 4 lanes under a device-memory budget
           vs a plain one-query-at-a-time server: queries/s, p50/p99 latency,
           and a sample of served results compared with the query run alone.
  cpu_baseline  the reference CPU BFS (cpuBFS::cpuBFS built from its
           sources) on one pinned host core.

Kernel configurations come from the persisted autotuner table
(gpu_graphs_amd/tuned.json) when it has an entry for the graph, else from the
autotuner run here; --retune always runs the autotuner, --tuning-out writes
what it chose to a table file.

Roofline (algorithmic bytes of one BFS from source s, int32 levels):
    B(s) = 4 n + 28 n_R + 8 m_R
  4 n level init; per reached vertex 16 B of offsets + 12 B of level write,
  enqueue and dequeue; per scanned CSR entry 4 B column + 4 B status probe.
`roofline.achieved` (= `alg_gbs`) divides the mean B(s) by the mean device
time of a BFS: from the start of its first kernel to the store of its final
control block by its last kernel, both stamped with the GPU's constant-rate
clock (wall_clock64) on the handle's stream (a start event recorded ahead of
the launches would keep the GPU idle for its own host time).  `roofline.traffic` is the HBM bytes per BFS that rocprofv3's
FETCH_SIZE and WRITE_SIZE counters gave for the same configuration
(profiles/bfs_rmat22_traffic.json), and `hbm_gbs` divides it by the same
device time: the bandwidth actually moved.  The direction-optimising search
reads fewer bytes than B(s) counts, so `alg_gbs` can exceed `hbm_gbs`.
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
SEED = 20260003          # single-source R-MAT s22, and the 7000 x 7000 grid
SEED_BATCH = 20260004    # batched R-MAT s24 (also the s24 graph of the served mix)
SEED_SERVE = 20260005    # the other served graphs and the query stream
RMAT_ABC = (0.57, 0.19, 0.19)
GRID_SIDE = 7000

WL_RMAT22 = f"rmat-s22-ef16-seed{SEED}"
WL_RMAT24 = f"rmat-s24-ef16-seed{SEED_BATCH}"
WL_GRID = f"grid-{GRID_SIDE}x{GRID_SIDE}-seed{SEED}"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# -- workloads (the parity tests build the same graphs and sources) ---------

def rmat22_graph(gg, scale=22, edge_factor=16):
    return gg.Graph.rmat(scale, edge_factor, *RMAT_ABC, seed=SEED, permute=True)


def rmat24_graph(gg):
    return gg.Graph.rmat(24, 16, *RMAT_ABC, seed=SEED_BATCH, permute=True)


def grid_graph(gg):
    return gg.Graph.grid(GRID_SIDE, GRID_SIDE, 0.05, seed=SEED)


def served_specs(gg, g24=None):
    """(name, constructor) of the 8 graphs of the served mix."""
    return [("rmat16", lambda: gg.Graph.rmat(16, 16, *RMAT_ABC, seed=SEED_SERVE)),
            ("rmat18", lambda: gg.Graph.rmat(18, 16, *RMAT_ABC, seed=SEED_SERVE)),
            ("rmat20", lambda: gg.Graph.rmat(20, 16, *RMAT_ABC, seed=SEED_SERVE)),
            ("rmat22", lambda: gg.Graph.rmat(22, 16, *RMAT_ABC, seed=SEED_SERVE)),
            ("rmat24", lambda: g24 if g24 is not None else rmat24_graph(gg)),
            ("grid1M", lambda: gg.Graph.grid(1000, 1000, 0.05, seed=SEED_SERVE)),
            ("grid4M", lambda: gg.Graph.grid(2000, 2000, 0.05, seed=SEED_SERVE)),
            ("grid16M", lambda: gg.Graph.grid(4000, 4000, 0.05, seed=SEED_SERVE))]


def pick_sources(offsets, k, seed):
    """k distinct sources drawn uniformly (seeded) from the vertices with edges."""
    deg = np.diff(offsets)
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, cand.size), replace=False).astype(np.int32)


def rmat22_sources(G):
    return pick_sources(G.offsets, 64, SEED)


def rmat24_sources(G):
    return pick_sources(G.offsets, 64, SEED_BATCH)


def grid_sources(G):
    return pick_sources(G.offsets, 16, SEED)


# -- helpers ---------------------------------------------------------------

def sync(gg):
    gg._ffi.check(gg.lib().gb_device_synchronize(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def configure(gg, dg, workload, kind, tune_sources, args):
    """Applies the table's configuration for (workload, kind), or runs the
    autotuner (always with --retune); returns the tuning record."""
    from gpu_graphs_amd import tuning
    if not args.retune:
        t = tuning.apply(dg, workload, kind)
        if t is not None:
            return dict(t, source="table")
    t0 = time.time()
    t = dg.autotune(tune_sources, reps=2) if kind == "single" else dg.batch_autotune(tune_sources, reps=2)
    log(f"autotune {workload} {kind} ({time.time() - t0:.1f}s): {t}")
    if args.tuning_out:
        tuning.record(workload, kind, dg.n, dg.m, t, path=args.tuning_out)
    return dict(t, source="autotune")


def traffic_profile(name, run_config):
    """HBM bytes per unit from a committed counter profile, when it was taken
    at exactly the configuration of this run."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, "no profile"
    with open(path) as f:
        prof = json.load(f)
    pc = prof.get("config", {})
    if not isinstance(pc, dict) or any(pc.get(k) != v for k, v in run_config.items()):
        return None, f"profile taken at {pc}, run at {run_config}"
    return prof, "profile taken at this run's configuration"


def cpu_baseline(offsets, cols, sources, gpu_levels_fn, budget_s):
    """Times the reference CPU BFS (or the oracle port) on one pinned core."""
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
        "cpu_model": cpu_model(),
        "ms_per_bfs": round(total_s * 1e3 / max(done, 1), 2),
        "parity_mismatches": mismatches,
    }


def explicit_kernel_flags(args):
    return any(x is not None for x in (args.block, args.k, args.alpha, args.beta, args.dense_edges)) or \
        args.no_bottom_up


# -- config 2: single source, R-MAT s22 ------------------------------------

def single_source(gg, args):
    t0 = time.time()
    G = rmat22_graph(gg, args.scale, args.edge_factor)
    n, m = G.n, G.m
    log(f"generated rmat s{args.scale} ef{args.edge_factor}: n={n} m={m} in {time.time() - t0:.1f}s")
    sources = rmat22_sources(G)
    dg = gg.DeviceGraph(G, 0)
    cfg = dg.get_config()
    if args.mode is not None:
        cfg["mode"] = args.mode
    if args.blocks_per_cu is not None:
        cfg["blocks_per_cu"] = args.blocks_per_cu
    dg.config(**cfg)
    if explicit_kernel_flags(args):
        cfg = dg.get_config()
        if args.block:
            cfg["block"] = args.block
        if args.k:
            cfg["edges_per_thread"] = args.k
        if args.dense_edges:
            cfg["dense_edges"] = args.dense_edges
        dg.config(**cfg)
        dg.direction_config(not args.no_bottom_up, args.alpha or 0, args.beta or 0)
        tuning = {"source": "flags"}
    else:
        workload = WL_RMAT22 if (args.scale, args.edge_factor) == (22, 16) else \
            f"rmat-s{args.scale}-ef{args.edge_factor}-seed{SEED}"
        tuning = configure(gg, dg, workload, "single", pick_sources(G.offsets, 8, SEED + 1), args)
    d_levels = gg.DeviceArray(0, np.int32, n)
    kcfg, dcfg = dg.get_config(), dg.get_direction_config()
    log(f"uploaded; config {kcfg} {dcfg}")

    for _ in range(args.warmup):
        for s in sources:
            dg.bfs_device(int(s), d_levels)
    sync(gg)
    stats = []
    srcs = [int(s) for s in sources]
    t_begin = time.perf_counter()
    for _ in range(args.steps):
        for s in srcs:
            stats.append(dg.bfs_device(s, d_levels, raw=True))
    sync(gg)
    wall = time.perf_counter() - t_begin
    stats = [st.as_dict() for st in stats]
    log(f"timed {args.steps} passes of {len(sources)} BFS in {wall * 1e3:.1f} ms")

    nb = len(stats)
    edges = sum(s["m_reached"] / 2 for s in stats)
    dev_s = sum(s["ms"] for s in stats) / 1e3
    b_alg = [4 * n + 28 * s["n_reached"] + 8 * s["m_reached"] for s in stats]
    kernel_ms = dev_s * 1e3 / nb
    alg_gbs = (sum(b_alg) / nb) / (kernel_ms / 1e3) / 1e9
    t_roof = sum(b / (HBM_PEAK_GBS * 1e9) for b in b_alg)
    # Per source: its mean device time over the passes.
    k = len(sources)
    per_gteps = []
    for i in range(k):
        mine = stats[i::k]
        per_gteps.append(mine[0]["m_reached"] / 2 / (sum(x["ms"] for x in mine) / len(mine) / 1e3) / 1e9)

    # The levels to the host (the reference's get_device_distance), outside
    # the timed region: one host-pointer search per source.
    copy_ms = [dg.bfs(int(s))[2]["copy_ms"] for s in sources[:16]]
    d2h_ms = float(np.mean(copy_ms))

    run_cfg = {"block": kcfg["block"], "edges_per_thread": kcfg["edges_per_thread"],
               "dense_edges": kcfg["dense_edges"], "alpha": dcfg["alpha"], "beta": dcfg["beta"]}
    prof, prof_note = traffic_profile(f"bfs_rmat{args.scale}_traffic.json", run_cfg)
    traffic = prof["hbm_bytes_per_bfs"] if prof else None
    hbm_gbs = traffic / (kernel_ms / 1e3) / 1e9 if traffic else None

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
            "workload": f"rmat-s{args.scale}-ef{args.edge_factor} single-source BFS; one step = one BFS from "
                        f"each of {k} seeded sources",
            "scale": args.scale, "edge_factor": args.edge_factor, "abc": list(RMAT_ABC),
            "seed": SEED, "n": n, "m_directed": m, "sources": k, "bfs_per_step": k,
            "kernel": kcfg, "direction": dcfg, "tuning": tuning,
        },
        "device_gteps": round(edges / dev_s / 1e9, 4),
        "gteps_per_source": {"harmonic_mean": round(k / sum(1 / g for g in per_gteps), 4),
                             "min": round(min(per_gteps), 4), "median": round(float(np.median(per_gteps)), 4)},
        "roofline_bfs_fraction": round(t_roof / dev_s, 4),
        "last_bottom_up_levels": list(dg.last_bottom_up()),
        "launch_hints": dict(zip(("depth", "bu_lo", "bu_hi"), dg.launch_hints())),
        "tail": dg.get_tail_config(),
        "levels_to_host": {"d2h_ms": round(d2h_ms, 4), "bytes": 4 * n,
                           "gbs": round(4 * n / (d2h_ms / 1e3) / 1e9, 2),
                           "gteps_incl_d2h": round(edges / (dev_s + nb * d2h_ms / 1e3) / 1e9, 4)},
        "roofline": {
            "bound": "hbm",
            "achieved": round(alg_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(alg_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_gbs": round(alg_gbs, 2),
            "hbm_gbs": round(hbm_gbs, 2) if hbm_gbs else None,
            "hbm_frac": round(hbm_gbs / HBM_PEAK_GBS, 4) if hbm_gbs else None,
            "traffic_note": prof_note,
            "achieved_note": "achieved = alg_gbs: B(s) = 4n + 28 n_R + 8 m_R bytes per search over its device time; "
                             "bottom-up levels read fewer bytes than B(s) charges, so frac can exceed 1. "
                             "hbm_gbs / hbm_frac divide the counter-measured traffic by the same time.",
            "kernel": "one BFS: its launch sequence on the handle's stream (bfs_clear_kernel, which seeds "
                      "the source, one bfs_expand_kernel per level, top-down or bottom-up, and bfs_tail_kernel "
                      "for the small last levels; bfs_compact_kernel only on directed graphs), timed from the "
                      "clear kernel's start to the final control-block store by the GPU's wall_clock64 stamps",
            "kernel_ms_avg": round(kernel_ms, 4),
            "alg_bytes_per_launch": int(sum(b_alg) / nb),
        },
        "cpu_baseline": None,
    }
    return result, G, dg, sources, d_levels


# -- config 3: the 7000 x 7000 road-like grid ------------------------------

def grid(gg, args):
    t0 = time.time()
    G = grid_graph(gg)
    n = G.n
    log(f"grid: {GRID_SIDE}x{GRID_SIDE} n={n} m={G.m} in {time.time() - t0:.1f}s")
    sources = grid_sources(G)
    dg = gg.DeviceGraph(G, 0)
    d_levels = gg.DeviceArray(0, np.int32, n)
    # Two searches from other seeded vertices teach the handle the graph's
    # depth and thin frontiers (one launch per level from then on).
    for s in pick_sources(G.offsets, 2, SEED + 7):
        dg.bfs_device(int(s), d_levels)
    sync(gg)
    stats = []
    t = time.perf_counter()
    for s in sources:
        stats.append(dg.bfs_device(int(s), d_levels))
    sync(gg)
    wall = time.perf_counter() - t
    edges = sum(s["m_reached"] / 2 for s in stats)
    dev_s = sum(s["ms"] for s in stats) / 1e3
    levels = [s["num_levels"] for s in stats]
    out = {
        "workload": f"{GRID_SIDE}x{GRID_SIDE} grid (spanning tree + 5% of the other grid edges), "
                    f"{len(sources)} seeded sources, one BFS each",
        "seed": SEED, "n": n, "m_directed": G.m, "sources": len(sources),
        "kernel": dg.get_config(), "deep_mode_used": dg.get_deep_config()["used"],
        "gteps": round(edges / wall / 1e9, 4),
        "device_gteps": round(edges / dev_s / 1e9, 4),
        "ms_per_bfs": round(wall * 1e3 / len(sources), 3),
        "device_ms_per_bfs": round(dev_s * 1e3 / len(sources), 3),
        "levels": {"mean": round(float(np.mean(levels)), 1), "min": int(min(levels)), "max": int(max(levels))},
        "us_per_level": round(dev_s * 1e6 / sum(levels), 3),
    }
    if not args.no_cpu_baseline:
        def gpu_levels(s):
            dg.bfs_device(s, d_levels)
            return d_levels.numpy()
        out["cpu_baseline"] = cpu_baseline(G.offsets, G.cols, sources, gpu_levels, args.grid_cpu_budget)
        out["parity_mismatches"] = out["cpu_baseline"]["parity_mismatches"]
    log(f"grid: {out}")
    return out


# -- config 4: one 64-source batch at R-MAT s24 ----------------------------

def batched(gg, args):
    """R-MAT s24: one 64-source batch vs 64 single-source BFS, same GPU.  Both
    paths run the configuration tuned for this graph."""
    t0 = time.time()
    G = rmat24_graph(gg)
    n = G.n
    log(f"batched: rmat s24 n={n} m={G.m} in {time.time() - t0:.1f}s")
    sources = rmat24_sources(G)
    dg = gg.DeviceGraph(G, 0)
    single_tuning = configure(gg, dg, WL_RMAT24, "single", pick_sources(G.offsets, 8, SEED_BATCH + 2), args)
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

    batch_tuning = configure(gg, dg, WL_RMAT24, "batch", pick_sources(G.offsets, 64, SEED_BATCH + 1), args)
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
    bcfg, bdir = dg.get_batch_config(), dg.get_batch_direction_config()
    run_cfg = {"block": bcfg["block"], "edges_per_thread": bcfg["edges_per_thread"],
               "sparse_edges": bcfg["sparse_edges"], "alpha": bdir["alpha"], "beta": bdir["beta"]}
    prof, prof_note = traffic_profile("msbfs_rmat24_traffic.json", run_cfg)
    out = {
        "workload": "rmat-s24-ef16, 64 seeded sources in one multi-source BFS",
        "seed": SEED_BATCH, "n": n, "m_directed": G.m, "sources": len(sources),
        "kernel": bcfg, "direction": bdir,
        "tuning": batch_tuning,
        "single_source_tuning": single_tuning,
        "batch_ms": round(batch_s * 1e3, 3),
        "batched_qps": round(len(sources) / batch_s, 1),
        "batched_gteps": round(edges / batch_s / 1e9, 3),
        "plain_qps": round(len(sources) / plain_s, 1),
        "plain_gteps": round(edges / plain_s / 1e9, 3),
        "speedup_vs_plain": round(plain_s / batch_s, 2),
        "num_levels": st["num_levels"],
        "rows_equal_single_source": len(sources) - mismatches,
        "parity_mismatches": mismatches,
        "traffic": prof["hbm_bytes_per_batch"] if prof else None,
        "hbm_gbs": round(prof["hbm_bytes_per_batch"] / batch_s / 1e9, 1) if prof else None,
        "traffic_note": prof_note,
    }
    log(f"batched: {out}")
    return out, G, dg


# -- config 5: served mix ---------------------------------------------------

def served(gg, args, g24=None, dg24=None):
    """8 resident graphs, Poisson query stream, batched server vs plain loop."""
    specs = served_specs(gg, g24)
    gs, dgs, cands = [], [], []
    for name, fn in specs:
        t = time.time()
        g = fn()
        gs.append(g)
        dgs.append(dg24 if (name == "rmat24" and dg24 is not None) else gg.DeviceGraph(g, 0))
        cands.append(np.nonzero(np.diff(g.offsets))[0].astype(np.int32))
        log(f"served: {name} n={g.n} m={g.m} ({time.time() - t:.1f}s)")
    checked = {"compared": 0, "mismatches": 0}

    def run(srv, gids, rate, nq, seed, sample_every=0):
        rng = np.random.default_rng(seed)
        gi = rng.integers(len(gs), size=nq)
        src = [int(rng.choice(cands[g])) for g in gi]
        gaps = rng.exponential(1.0 / rate, size=nq) if rate else np.zeros(nq)
        tickets = [None] * nq
        failures = []
        kept = []  # (graph, source, served levels) of the sampled queries
        lat_by_graph = [[] for _ in gs]  # total latency (ms) of each completed query, by graph

        def releaser():
            # Results are consumed and released as they complete, in any order.
            pending, nxt_i, finished = [], 0, 0
            while finished < nq:
                while nxt_i < nq and tickets[nxt_i] is not None:
                    pending.append((nxt_i, tickets[nxt_i]))
                    nxt_i += 1
                still = []
                for i, t in pending:
                    out = None
                    if sample_every and i % sample_every == 0:
                        out = np.empty(gs[gi[i]].n, dtype=np.int32)
                    rc, lat = srv.poll(t, out)
                    if rc == gg.GB_EAGAIN:
                        still.append((i, t))
                        continue
                    if rc != 0:
                        failures.append(rc)
                    else:
                        lat_by_graph[gi[i]].append((lat["total_ms"], lat["queue_ms"], lat["run_ms"]))
                        if out is not None:
                            kept.append((int(gi[i]), src[i], out))
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
        for g, s, lv in kept:  # each sampled result against the same query alone
            checked["compared"] += 1
            if not np.array_equal(lv, dgs[g].bfs(s)[0]):
                checked["mismatches"] += 1
        def pct(v, k, q):
            return round(float(np.percentile([x[k] for x in v], q)), 2)
        # total = queue (arrival -> batch start) + run (batch start -> done)
        per_graph = {specs[g][0]: {"queries": len(v), "p50_ms": pct(v, 0, 50), "p99_ms": pct(v, 0, 99),
                                   "queue_p99_ms": pct(v, 1, 99), "run_p99_ms": pct(v, 2, 99)}
                     for g, v in enumerate(lat_by_graph) if v}
        return {"offered_qps": round(rate, 1) if rate else "saturated", "queries": nq,
                "qps": round(st["qps"], 1), "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                "mean_batch": round(st["mean_batch"], 2), "failed": len(failures),
                "memory_used": st["memory_used"], "workspace_drops": st["workspace_drops"], "per_graph": per_graph}

    out = {"graphs": [{"name": s[0], "n": g.n, "m_directed": g.m} for s, g in zip(specs, gs)],
           "seed": SEED_SERVE, "memory_budget": args.serve_budget}
    plain = gg.Server(lanes=1, plain=True)
    pids = [plain.add_graph(d) for d in dgs]
    run(plain, pids, 0, 16, 1)
    p_sat = run(plain, pids, 0, args.serve_plain_queries, 2, sample_every=8)
    log(f"served plain saturated: {p_sat}")
    batch = gg.Server(lanes=4, max_batch=64, memory_budget=args.serve_budget)
    bids = [batch.add_graph(d) for d in dgs]
    run(batch, bids, 0, 64, 3)
    b_sat = run(batch, bids, 0, args.serve_batch_queries, 4, sample_every=64)
    log(f"served batched saturated: {b_sat}")
    sweep = []
    cap = p_sat["qps"]
    for f in (0.5, 1.0, 2.0, 4.0):
        rate = cap * f
        nq = int(max(rate * args.serve_seconds, 24))
        row = {"rate_x_plain": f, "batched": run(batch, bids, rate, nq, 10 + int(f * 4), sample_every=16)}
        if f <= 1.0:
            row["plain"] = run(plain, pids, rate, nq, 10 + int(f * 4))
        log(f"served sweep: {row}")
        sweep.append(row)
    plain.close()
    batch.close()
    out.update({"plain_saturated": p_sat, "batched_saturated": b_sat, "sweep": sweep,
                "saturated_speedup_vs_plain": round(b_sat["qps"] / max(p_sat["qps"], 1e-9), 2),
                "results_compared_with_alone": checked["compared"],
                "parity_mismatches": checked["mismatches"]})
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
    ap.add_argument("--steps", type=int, default=20, help="timed passes over the 64 sources")
    ap.add_argument("--warmup", type=int, default=3, help="untimed passes over the 64 sources")
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
    ap.add_argument("--retune", action="store_true", help="run the autotuners instead of applying the table")
    ap.add_argument("--tuning-out", default=None, help="write what the autotuners chose to this table file")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work (R-MAT s22)")
    ap.add_argument("--grid-cpu-budget", type=float, default=8.0, help="seconds of CPU baseline work (grid)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--serve-budget", type=int, default=128 << 30,
                    help="device bytes for the batched server's workspaces and result buffers")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="arrival window of each swept rate")
    ap.add_argument("--serve-plain-queries", type=int, default=96)
    ap.add_argument("--serve-batch-queries", type=int, default=2048)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (byte-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import gpu_graphs_amd as gg
    assert gg.device_count() > 0, "no GPU visible"

    result = {"metric": "GTEPS single-source BFS + batched queries/s & p99 latency vs plain, same GPU"}
    if not args.no_single:
        result, G, dg, sources, d_levels = single_source(gg, args)
        if not args.no_cpu_baseline:
            def gpu_levels(s):
                dg.bfs_device(s, d_levels)
                return d_levels.numpy()
            result["cpu_baseline"] = cpu_baseline(G.offsets, G.cols, sources, gpu_levels, args.cpu_budget)
        del dg, d_levels, G
    if not args.no_grid:
        result["grid"] = grid(gg, args)
    g24 = dg24 = None
    if not args.no_batched:
        result["batched"], g24, dg24 = batched(gg, args)
    if not args.no_served:
        result["served"] = served(gg, args, g24, dg24)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
