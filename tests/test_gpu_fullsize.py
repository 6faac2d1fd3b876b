"""Parity at the full sizes of the workloads: R-MAT scale 22 (single
source, 4.2M vertices), the 7000 x 7000 road-like grid (49M vertices, deep
mode once learned) and the R-MAT scale 24 batch of 64 sources (16.8M
vertices).  Levels are bit-exact against the CPU oracle (or, for the batch,
against single-source searches, which are themselves checked against the
oracle), parents pass the BFS-tree rules, and the statistics match."""
import numpy as np
import pytest

import gpu_graphs_amd as gg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RMAT_ABC = (0.57, 0.19, 0.19)


def log(msg):
    print(msg, flush=True)


def sources_of(G, k, seed):
    deg = np.diff(G.offsets)
    return [int(s) for s in np.random.default_rng(seed).choice(np.nonzero(deg)[0], k, replace=False)]


def check_search(G, dg, s):
    lv, par, st = dg.bfs(s, parents=True)
    want, _, ost = O.bfs(G.offsets, G.cols, s)
    np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
    assert O.check_parents(G.offsets, G.cols, s, lv, par) == 0, s
    assert (st["n_reached"], st["m_reached"], st["num_levels"]) == (
        ost["n_reached"], ost["m_reached"], ost["num_levels"]), s
    return st


def test_rmat22_sampled_sources(gpu):
    G = gg.Graph.rmat(22, 16, *RMAT_ABC, seed=20260003, permute=True)
    dg = gg.DeviceGraph(G, gpu)
    for s in sources_of(G, 5, 20260003):
        log(f"rmat22 source {s}")
        check_search(G, dg, s)
    assert dg.last_bottom_up()[0] >= 1  # the direction-optimised path ran


def test_road_grid_7000(gpu):
    G = gg.Graph.grid(7000, 7000, 0.05, seed=20260003)
    assert G.n == 49_000_000
    dg = gg.DeviceGraph(G, gpu)
    rng = np.random.default_rng(20260003)
    used = []
    for s in [int(x) for x in rng.choice(G.n, 4, replace=False)]:
        log(f"grid7000 source {s}")
        st = check_search(G, dg, s)
        assert st["num_levels"] > 1000
        used.append(dg.get_deep_config()["used"])
    assert used[-1], used  # deep mode engaged after the first searches


def test_rmat24_batch_of_64(gpu):
    G = gg.Graph.rmat(24, 16, *RMAT_ABC, seed=20260004, permute=True)
    dg = gg.DeviceGraph(G, gpu)
    srcs = np.array(sources_of(G, 64, 20260004), dtype=np.int32)
    n = G.n
    rows = gg.DeviceArray(gpu, np.int32, 64 * n)
    one = gg.DeviceArray(gpu, np.int32, n)
    log("rmat24 batch of 64")
    dg.bfs_batch_device(srcs, rows)
    for i, s in enumerate(srcs):
        dg.bfs_device(int(s), one)
        np.testing.assert_array_equal(rows.numpy_slice(i * n, n), one.numpy(), err_msg=f"row {i} source {s}")
    for i in (0, 63):
        want = O.bfs(G.offsets, G.cols, int(srcs[i]))[0]
        np.testing.assert_array_equal(rows.numpy_slice(i * n, n), want, err_msg=f"row {i} vs oracle")
