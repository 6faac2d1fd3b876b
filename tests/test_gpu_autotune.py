"""The autotuners (gb_bfs_autotune, gb_batch_autotune): they time real
configurations on the device, leave a supported configuration on the handle,
and every search after tuning is still bit-exact with the CPU oracle.  A
tuning result applied to a fresh handle reproduces the same configuration."""
import numpy as np
import pytest

import gpu_graphs_amd as gg
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_single_source_autotune_keeps_parity(gpu):
    G = gg.Graph.rmat(15, seed=91)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    sources = np.nonzero(deg)[0][:4].astype(np.int32)
    log("autotune rmat15, 4 sources")
    res = dg.autotune(sources, reps=2)
    assert res["tried"] >= 8
    assert 0 < res["best_ms"] <= res["default_ms"]
    cfg = dg.get_config()
    assert (cfg["block"], cfg["edges_per_thread"]) == (res["block"], res["edges_per_thread"])
    assert cfg["dense_edges"] == res["frontier_edges"]
    d = dg.get_direction_config()
    assert (d["alpha"], d["beta"]) == (res["alpha"], res["beta"])
    for s in list(sources) + [int(np.argmax(deg))]:
        levels, parents, _ = dg.bfs(int(s), parents=True)
        want = O.bfs(G.offsets, G.cols, int(s))[0]
        np.testing.assert_array_equal(levels, want)
        assert O.check_parents(G.offsets, G.cols, int(s), levels, parents) == 0

    fresh = gg.DeviceGraph(G, gpu)
    fresh.apply_tuning(res)
    assert fresh.get_config()["block"] == res["block"]
    assert fresh.get_config()["dense_edges"] == res["frontier_edges"]
    levels, _, _ = fresh.bfs(int(sources[0]))
    np.testing.assert_array_equal(levels, O.bfs(G.offsets, G.cols, int(sources[0]))[0])


def test_batch_autotune_keeps_parity(gpu):
    G = gg.Graph.rmat(14, seed=92)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(5)
    sources = rng.choice(np.nonzero(deg)[0], 64, replace=False).astype(np.int32)
    log("batch autotune rmat14, 64 sources")
    res = dg.batch_autotune(sources, reps=2)
    assert res["tried"] >= 8
    assert 0 < res["best_ms"] <= res["default_ms"]
    cfg = dg.get_batch_config()
    assert (cfg["block"], cfg["edges_per_thread"], cfg["sparse_edges"]) == (
        res["block"], res["edges_per_thread"], res["frontier_edges"])
    assert dg.get_batch_direction_config()["alpha"] == res["alpha"]
    want = O.bfs_multi(G.offsets, G.cols, sources)
    levels, _ = dg.bfs_batch(sources)
    np.testing.assert_array_equal(levels, want)
    other = rng.choice(G.n, 17).astype(np.int32)
    levels, _ = dg.bfs_batch(other)
    np.testing.assert_array_equal(levels, O.bfs_multi(G.offsets, G.cols, other))

    fresh = gg.DeviceGraph(G, gpu)
    fresh.apply_batch_tuning(res)
    levels, _ = fresh.bfs_batch(sources)
    np.testing.assert_array_equal(levels, want)


def test_autotune_on_a_directed_graph_skips_direction_knobs(gpu):
    n = 3000
    rng = np.random.default_rng(8)
    src = rng.integers(0, n, 20000)
    dst = rng.integers(0, n, 20000)
    G = gg.Graph.from_edges(n, src, dst)
    assert not G.is_symmetric
    dg = gg.DeviceGraph(G, gpu)
    res = dg.autotune([0, 1], reps=1)
    assert res["tried"] == 1 + 5 + 2  # shapes and the frontier threshold only
    levels, _, _ = dg.bfs(0)
    np.testing.assert_array_equal(levels, O.bfs(G.offsets, G.cols, 0)[0])


def test_autotune_rejects_bad_sources(gpu):
    G = gg.Graph.rmat(10, seed=1)
    dg = gg.DeviceGraph(G, gpu)
    with pytest.raises(gg.GBError) as e:
        dg.autotune([G.n])
    assert e.value.code == gg.GB_ERANGE
    with pytest.raises(gg.GBError) as e:
        dg.batch_autotune(np.zeros(65, np.int32))
    assert e.value.code == gg.GB_EINVAL
