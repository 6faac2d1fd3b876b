"""Parity at the full sizes of the benchmark workloads, for every source the
benchmark times, under the configuration it runs (the persisted autotuner
table, gpu_graphs_amd/tuned.json):

* R-MAT scale 22 (4.2M vertices), all 64 sources of the single-source bench;
* the 7000 x 7000 road-like grid (49M vertices, deep mode once learned), all
  16 sources of the grid bench;
* two different 64-source batches at R-MAT scale 24 (16.8M vertices) on one
  handle, every row of both: the second batch runs on level lines the first
  one left behind (they are never cleared).

Levels are bit-exact against the CPU oracle (the restatement of the
reference's cpuBFS::cpuBFS, 1D Partition/src/cpuBFS.cpp:5-62), parents pass
the BFS-tree rules and the statistics match.  The oracle runs one source per
host thread (the ctypes calls release the GIL)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bench
import gpu_graphs_amd as gg
from gpu_graphs_amd import tuning
from oracle import oracle as O

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8))


def log(msg):
    print(msg, flush=True)


def oracle_task(off, cols, s, lv, par):
    """Oracle levels and statistics for s, and the BFS-tree check of the GPU's parents."""
    want, _, ost = O.bfs(off, cols, s)
    return want, ost, O.check_parents(off, cols, s, lv, par)


def check_every_source(G, dg, sources, min_levels=0):
    off, cols = G.offsets, G.cols
    with ThreadPoolExecutor(WORKERS) as pool:
        futs = []
        for s in sources:
            lv, par, st = dg.bfs(int(s), parents=True)
            futs.append((int(s), lv, st, pool.submit(oracle_task, off, cols, int(s), lv, par)))
        for s, lv, st, f in futs:
            want, ost, tree = f.result()
            np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
            assert tree == 0, (s, tree)
            assert (st["n_reached"], st["m_reached"], st["num_levels"]) == (
                ost["n_reached"], ost["m_reached"], ost["num_levels"]), s
            assert st["num_levels"] >= min_levels, (s, st["num_levels"])
    log(f"{len(sources)} sources equal the oracle")


def test_rmat22_every_bench_source(gpu):
    G = bench.rmat22_graph(gg)
    dg = gg.DeviceGraph(G, gpu)
    log(f"rmat22 tuning {tuning.apply(dg, bench.WL_RMAT22, 'single')}")
    sources = bench.rmat22_sources(G)
    assert len(sources) == 64
    check_every_source(G, dg, sources)
    assert dg.last_bottom_up()[0] >= 1  # the direction-optimised path ran


def test_road_grid_7000_every_bench_source(gpu):
    G = bench.grid_graph(gg)
    assert G.n == 49_000_000
    dg = gg.DeviceGraph(G, gpu)
    sources = bench.grid_sources(G)
    assert len(sources) == 16
    # Two thin, deep searches teach the handle the graph's shape; the third
    # runs one launch per level (deep mode).
    check_every_source(G, dg, sources[:3], min_levels=1000)
    assert dg.get_deep_config()["used"]
    check_every_source(G, dg, sources[3:], min_levels=1000)


def test_rmat24_two_batches_every_row(gpu):
    G = bench.rmat24_graph(gg)
    dg = gg.DeviceGraph(G, gpu)
    log(f"rmat24 batch tuning {tuning.apply(dg, bench.WL_RMAT24, 'batch')}")
    n = G.n
    off, cols = G.offsets, G.cols
    first = bench.rmat24_sources(G)
    second = bench.pick_sources(G.offsets, 64, bench.SEED_BATCH + 17)
    assert len(set(first.tolist()) & set(second.tolist())) < 8
    rows = gg.DeviceArray(gpu, np.int32, 64 * n)
    with ThreadPoolExecutor(WORKERS) as pool:
        for name, srcs in (("first", first), ("second", second)):
            futs = [pool.submit(O.bfs, off, cols, int(s)) for s in srcs]
            log(f"rmat24 {name} batch of 64")
            st = dg.bfs_batch_device(srcs, rows, stats=True)
            for i, s in enumerate(srcs):
                want, _, ost = futs[i].result()
                np.testing.assert_array_equal(rows.numpy_slice(i * n, n), want,
                                              err_msg=f"{name} batch row {i} source {s}")
                assert (st["n_reached"][i], st["m_reached"][i]) == (ost["n_reached"], ost["m_reached"]), (name, i)
                futs[i] = None
            log(f"rmat24 {name} batch: 64 rows equal the oracle")
