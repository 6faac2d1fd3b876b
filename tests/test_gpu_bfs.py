"""GPU parity of single-source BFS (gb_bfs / gb_bfs_device) against the oracle.

Levels must be bit-exact: against the scipy goldens for the reference's own
fixture graphs (every source), against the CPU oracle for generated graphs,
and, at full size, against the oracle for sampled sources plus the
size-independent level invariants.  Parents must pass the Graph500 rules.
"""
import os

import numpy as np
import pytest

import gpu_graphs_amd as gg
from conftest import FIXTURES, GOLDEN, load_fixture
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CONFIGS = [(256, 1), (256, 2), (256, 4), (256, 8), (512, 2), (512, 4), (1024, 1), (1024, 2), (128, 8), (64, 16)]


def log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", list(FIXTURES))
def test_fixture_levels_every_source(gpu, name, golden_levels):
    g = load_fixture(name)
    sources = golden_levels[f"{name}__sources"]
    want = golden_levels[f"{name}__levels"]
    dg = g.device_graph(gpu)
    log(f"fixture {name} n={g.num_nodes} m={g.num_edges} sources={len(sources)}")
    for i, s in enumerate(sources):
        lv, par, st = dg.bfs(int(s), parents=True)
        np.testing.assert_array_equal(lv, want[i], err_msg=f"{name} source {s}")
        assert O.check_parents(g.row_offset, g.col_idx, int(s), lv, par) == 0, (name, s)
        assert st["n_reached"] == int((want[i] >= 0).sum())
        assert st["num_levels"] == int(want[i].max()) + 1


def test_reference_api_mirror(gpu):
    g = gg.csr(os.path.join(GOLDEN, "graphs", "tree.txt"))
    b = gg.gpuBFS(g, 1)
    assert b.host_distance.tolist() == [-1, 0, 1, 1, 2, 2, 2, 2]
    assert b.iteration == 3 and b.total_edges_traversed == 6 and b.exec_time > 0


def test_er_workload_golden(gpu):
    g = gg.csr(os.path.join(GOLDEN, "er_10200.txt"))
    want = np.load(os.path.join(GOLDEN, "er_10200_levels.npy"))
    lv, _, st = g.device_graph(gpu).bfs(0)
    np.testing.assert_array_equal(lv, want)
    assert st["n_reached"] == 10000


# dense_edges: 1 = every level rebuilds its queue by bitmap compaction,
# 2**32-1 = every level queues by atomics, 0 = the default threshold.
DENSE = [1, 2**32 - 1, 0]


@pytest.mark.parametrize("dense", DENSE)
@pytest.mark.parametrize("mode", [gg.GB_MODE_GRAPH, gg.GB_MODE_ASYNC, gg.GB_MODE_LEVEL_LOOP])
@pytest.mark.parametrize("block,k", CONFIGS)
def test_rmat_all_kernel_shapes(gpu, mode, block, k, dense):
    G = gg.Graph.rmat(14, seed=41)
    dg = gg.DeviceGraph(G, gpu)
    dg.config(block=block, edges_per_thread=k, mode=mode, dense_edges=dense)
    off, cols = G.offsets, G.cols
    deg = np.diff(off)
    rng = np.random.default_rng(block * 10 + k + mode)
    log(f"rmat14 block={block} k={k} mode={mode} dense={dense}")
    for s in rng.choice(np.nonzero(deg)[0], 4, replace=False):
        lv, par, st = dg.bfs(int(s), parents=True)
        want, _, ost = O.bfs(off, cols, int(s))
        np.testing.assert_array_equal(lv, want)
        assert O.check_parents(off, cols, int(s), lv, par) == 0
        assert st["n_reached"] == ost["n_reached"] and st["m_reached"] == ost["m_reached"]
        assert st["num_levels"] == ost["num_levels"]


@pytest.mark.parametrize("dense", [1, 5000])
@pytest.mark.parametrize("name", ["wiki-Vote", "test2", "test_isolated_node", "test4"])
def test_fixtures_through_compaction(gpu, name, dense, golden_levels):
    # Directed fixtures with sinks: degree-0 vertices are reached but never
    # queued, which the compaction path counts separately.
    g = load_fixture(name)
    dg = g.device_graph(gpu)
    dg.config(dense_edges=dense)
    sources = golden_levels[f"{name}__sources"]
    want = golden_levels[f"{name}__levels"]
    for i, s in enumerate(sources[:16]):
        lv, par, st = dg.bfs(int(s), parents=True)
        np.testing.assert_array_equal(lv, want[i], err_msg=f"{name} source {s}")
        assert O.check_parents(g.row_offset, g.col_idx, int(s), lv, par) == 0
        assert st["n_reached"] == int((want[i] >= 0).sum())
        assert st["num_levels"] == int(want[i].max()) + 1
    dg.config()


@pytest.mark.parametrize("scale,seed", [(16, 1), (18, 2)])
def test_rmat_levels_match_oracle(gpu, scale, seed):
    G = gg.Graph.rmat(scale, seed=seed)
    dg = gg.DeviceGraph(G, gpu)
    off, cols = G.offsets, G.cols
    deg = np.diff(off)
    rng = np.random.default_rng(seed)
    srcs = list(rng.choice(np.nonzero(deg)[0], 8, replace=False)) + [int(np.argmax(deg))]
    log(f"rmat{scale} n={G.n} m={G.m}")
    for s in srcs:
        lv, _, st = dg.bfs(int(s))
        want, _, ost = O.bfs(off, cols, int(s))
        np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
        assert st["m_reached"] == ost["m_reached"]


def test_concurrent_handles_on_large_graph(gpu):
    # Four threads, four handles of one 8.4M-vertex graph (2048 compaction
    # chunks), BFS at the same time on four streams: the level kernels share
    # the CUs, and two handles rebuild every level's queue by compaction.
    import threading

    G = gg.Graph.rmat(23, edge_factor=2, seed=21)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(21)
    srcs = [int(s) for s in rng.choice(np.nonzero(deg)[0], 3, replace=False)]
    want = [O.bfs(G.offsets, G.cols, s)[0] for s in srcs]
    dgs = [gg.DeviceGraph(G, gpu) for _ in range(4)]
    dgs[0].config(dense_edges=1)
    dgs[1].config(dense_edges=1)
    log(f"rmat23 ef2 n={G.n} m={G.m}, 4 threads x {len(srcs)} sources")
    results, errors = {}, []

    def worker(i):
        try:
            for rep in range(2):
                for j, s in enumerate(srcs):
                    results[(i, rep, j)] = dgs[i].bfs(s)[0]
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert len(results) == 4 * 2 * len(srcs)
    for (i, rep, j), lv in results.items():
        np.testing.assert_array_equal(lv, want[j], err_msg=f"handle {i} rep {rep} source {srcs[j]}")


# Direction optimisation: (alpha, beta) = defaults, bottom-up as early as
# possible, back top-down early, bottom-up early and never back, and rarely
# bottom-up.
@pytest.mark.parametrize("alpha,beta", [(14, 24), (1000, 24), (14, 2), (1000, 100000), (2, 1000)])
@pytest.mark.parametrize("scale,seed", [(14, 71), (17, 72)])
def test_direction_optimising_matches_oracle(gpu, scale, seed, alpha, beta):
    G = gg.Graph.rmat(scale, seed=seed)
    dg = gg.DeviceGraph(G, gpu)
    dg.direction_config(True, alpha, beta)
    cfg = dg.get_direction_config()
    assert cfg["symmetric"] and cfg["alpha"] == alpha and cfg["beta"] == beta
    off, cols = G.offsets, G.cols
    deg = np.diff(off)
    rng = np.random.default_rng(seed + alpha)
    srcs = [int(s) for s in rng.choice(np.nonzero(deg)[0], 6, replace=False)] + [int(np.argmax(deg))]
    srcs.append(int(np.nonzero(deg == 0)[0][0]))  # isolated source
    log(f"rmat{scale} alpha={alpha} beta={beta}")
    bottom_up = []
    for s in srcs:
        lv, par, st = dg.bfs(s, parents=True)
        want, _, ost = O.bfs(off, cols, s)
        np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
        assert O.check_parents(off, cols, s, lv, par) == 0, s
        assert st["n_reached"] == ost["n_reached"] and st["m_reached"] == ost["m_reached"]
        assert st["num_levels"] == ost["num_levels"]
        bottom_up.append(dg.last_bottom_up())
    if alpha >= 14:
        assert any(f >= 1 for f, _ in bottom_up), bottom_up  # bottom-up levels did run


def test_direction_optimising_on_off_and_modes(gpu):
    G = gg.Graph.rmat(18, seed=73)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    srcs = [int(s) for s in np.random.default_rng(73).choice(np.nonzero(deg)[0], 4, replace=False)]
    want = {s: O.bfs(G.offsets, G.cols, s)[0] for s in srcs}
    for enable in (True, False):
        dg.direction_config(enable)
        for mode in (gg.GB_MODE_GRAPH, gg.GB_MODE_ASYNC, gg.GB_MODE_LEVEL_LOOP):
            dg.config(mode=mode)
            for s in srcs:
                lv, par, _ = dg.bfs(s, parents=True)
                np.testing.assert_array_equal(lv, want[s], err_msg=f"enable={enable} mode={mode} source {s}")
                assert O.check_parents(G.offsets, G.cols, s, lv, par) == 0
                if not enable:
                    assert dg.last_bottom_up() == (-1, -1)


@pytest.mark.parametrize("block,k", CONFIGS)
def test_direction_optimising_kernel_shapes_dense_every_level(gpu, block, k):
    # dense_edges=1: every top-down level goes through compaction and leaves a
    # frontier bitmap, so bottom-up can start at any level
    G = gg.Graph.rmat(14, seed=74)
    dg = gg.DeviceGraph(G, gpu)
    dg.config(block=block, edges_per_thread=k, dense_edges=1)
    deg = np.diff(G.offsets)
    for s in np.random.default_rng(block * k).choice(np.nonzero(deg)[0], 3, replace=False):
        lv, _, _ = dg.bfs(int(s))
        np.testing.assert_array_equal(lv, O.bfs(G.offsets, G.cols, int(s))[0])


def test_bottom_up_then_back_top_down(gpu):
    # An R-MAT core joined by one edge to a 300 x 60 grid: the core's big
    # levels go bottom-up, then the thin, shrinking grid frontiers (far below
    # n / beta vertices) send the search back top-down (bitmap turned into a
    # queue).
    core = gg.Graph.rmat(14, seed=77)
    grid = gg.Graph.grid(300, 60, 0.05, seed=78)
    nc = core.n
    cs = np.repeat(np.arange(nc), np.diff(core.offsets)).astype(np.int32)
    gs = (np.repeat(np.arange(grid.n), np.diff(grid.offsets)) + nc).astype(np.int32)
    hub = int(np.argmax(np.diff(core.offsets)))
    src = np.concatenate([cs, gs, [hub, nc]]).astype(np.int32)
    dst = np.concatenate([core.cols, grid.cols + nc, [nc, hub]]).astype(np.int32)
    G = gg.Graph.from_edges(nc + grid.n, src, dst)
    assert G.is_symmetric
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    srcs = [hub] + [int(s) for s in np.random.default_rng(77).choice(np.nonzero(deg[:nc])[0], 3, replace=False)]
    switched = False
    for s in srcs:
        lv, par, st = dg.bfs(s, parents=True)
        want, _, ost = O.bfs(G.offsets, G.cols, s)
        np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
        assert O.check_parents(G.offsets, G.cols, s, lv, par) == 0
        assert st["n_reached"] == ost["n_reached"] and st["m_reached"] == ost["m_reached"]
        first, last = dg.last_bottom_up()
        if first >= 1 and last + 2 < st["num_levels"]:
            switched = True  # bottom-up levels, then top-down to the end
    assert switched


def test_directed_graph_stays_top_down(gpu):
    # a directed graph must not be searched bottom-up (in-edges != out-edges)
    rng = np.random.default_rng(75)
    n = 20000
    src = rng.integers(0, n, 200000).astype(np.int32)
    dst = rng.integers(0, n, 200000).astype(np.int32)
    G = gg.Graph.from_edges(n, src, dst)
    assert not G.is_symmetric
    dg = gg.DeviceGraph(G, gpu)
    assert not dg.get_direction_config()["symmetric"]
    for s in (0, 17, 999):
        lv, _, _ = dg.bfs(s)
        np.testing.assert_array_equal(lv, O.bfs(G.offsets, G.cols, s)[0])
        assert dg.last_bottom_up() == (-1, -1)


def test_deep_grid_with_direction_optimisation(gpu):
    # thin frontiers: no bottom-up level; the learned window turns the
    # bottom-up launches off after the first search
    G = gg.Graph.grid(400, 300, 0.05, seed=76)
    dg = gg.DeviceGraph(G, gpu)
    for s in (0, 150 * 400 + 200, G.n - 1):
        lv, _, st = dg.bfs(s)
        want, _, ost = O.bfs(G.offsets, G.cols, s)
        np.testing.assert_array_equal(lv, want)
        assert st["num_levels"] == ost["num_levels"]


def check_search(G, dg, s, parents=True):
    lv, par, st = dg.bfs(s, parents=parents)
    want, _, ost = O.bfs(G.offsets, G.cols, s)
    np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
    if parents:
        assert O.check_parents(G.offsets, G.cols, s, lv, par) == 0, s
    assert st["n_reached"] == ost["n_reached"] and st["m_reached"] == ost["m_reached"], s
    assert st["num_levels"] == ost["num_levels"], s


@pytest.mark.parametrize("mode", [gg.GB_MODE_GRAPH, gg.GB_MODE_ASYNC, gg.GB_MODE_LEVEL_LOOP])
@pytest.mark.parametrize("block,k", [(256, 4), (1024, 1), (64, 16)])
def test_deep_mode_one_launch_per_level(gpu, mode, block, k):
    # Deep mode: every level in one launch of the deep kernel (a thread per
    # frontier vertex), block 0 of each level doing the bookkeeping.
    G = gg.Graph.grid(400, 300, 0.05, seed=86)
    dg = gg.DeviceGraph(G, gpu)
    dg.config(block=block, edges_per_thread=k, mode=mode)
    dg.deep_config(gg.GB_DEEP_ON)
    log(f"deep grid block={block} k={k} mode={mode}")
    for s in (0, 150 * 400 + 200, G.n - 1, 12345):
        check_search(G, dg, s)
        assert dg.get_deep_config() == {"mode": gg.GB_DEEP_ON, "used": True}


def test_deep_mode_on_wide_levels(gpu):
    # Forced on a graph with wide levels (a 3000-vertex path ending in an
    # R-MAT core): every level, hubs included, goes through the queue.
    core = gg.Graph.rmat(14, seed=87)
    nc, plen = core.n, 3000
    cs = np.repeat(np.arange(nc), np.diff(core.offsets))
    path = np.arange(nc, nc + plen)
    hub = int(np.argmax(np.diff(core.offsets)))
    src = np.concatenate([cs, path[:-1], path[1:], [path[-1], hub]]).astype(np.int32)
    dst = np.concatenate([core.cols, path[1:], path[:-1], [hub, path[-1]]]).astype(np.int32)
    G = gg.Graph.from_edges(nc + plen, src, dst)
    dg = gg.DeviceGraph(G, gpu)
    dg.deep_config(gg.GB_DEEP_ON)
    for s in (nc, nc + 1500, hub):
        check_search(G, dg, s)
    dg.deep_config(gg.GB_DEEP_OFF)
    check_search(G, dg, nc + 7)
    assert not dg.get_deep_config()["used"]


def test_deep_mode_row_lengths_and_edgeless_sources(gpu):
    # A 2000-vertex path whose every 100th vertex also has a fan of leaves,
    # with fans of 1..8 (claimed by the thread) and 9, 63, 64, 65, 300
    # entries (claimed by the wave 64 at a time), and some edgeless vertices.
    plen = 2000
    fans = [1, 2, 7, 8, 9, 63, 64, 65, 300] * 3
    src, dst, nxt = [], [], plen
    for i in range(plen - 1):
        src += [i, i + 1]
        dst += [i + 1, i]
    for j, f in enumerate(fans):
        hub = 100 * (j % 20) + 50
        for _ in range(f):
            src += [hub, nxt]
            dst += [nxt, hub]
            nxt += 1
    n = nxt + 5  # five edgeless vertices at the end
    G = gg.Graph.from_edges(n, np.array(src, dtype=np.int32), np.array(dst, dtype=np.int32))
    dg = gg.DeviceGraph(G, gpu)
    dg.deep_config(gg.GB_DEEP_ON)
    for s in (0, 1234, plen - 1, plen + 10, n - 1, n - 3):
        check_search(G, dg, s)
        assert dg.get_deep_config()["used"]


def test_deep_mode_auto_and_directed(gpu):
    # GB_DEEP_AUTO engages after deep, thin searches; R-MAT searches never
    # use it; directed graphs work the same.
    G = gg.Graph.grid(300, 250, 0.05, seed=88)
    dg = gg.DeviceGraph(G, gpu)
    assert dg.get_deep_config()["mode"] == gg.GB_DEEP_AUTO
    used = []
    for s in (0, 7, G.n // 2, G.n - 1, 999):
        check_search(G, dg, s, parents=False)
        used.append(dg.get_deep_config()["used"])
    assert not used[0] and used[-1], used
    R = gg.Graph.rmat(14, seed=89)
    dr = gg.DeviceGraph(R, gpu)
    for s in np.nonzero(np.diff(R.offsets))[0][:4]:
        check_search(R, dr, int(s), parents=False)
        assert not dr.get_deep_config()["used"]
    n = 5000
    src = np.arange(n - 1, dtype=np.int32)
    G2 = gg.Graph.from_edges(n, src, src + 1)  # one directed path
    dg2 = gg.DeviceGraph(G2, gpu)
    dg2.deep_config(gg.GB_DEEP_ON)
    for s in (0, 2500, n - 1):
        check_search(G2, dg2, s)


def test_graph_replay_serves_any_source_and_buffer(gpu):
    # The captured launch sequence reads the source and the output pointers
    # from device memory: replays with different sources, two level buffers,
    # with and without parents, interleaved with direct-launch searches.
    G = gg.Graph.rmat(15, seed=79)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    srcs = [int(s) for s in np.random.default_rng(79).choice(np.nonzero(deg)[0], 12, replace=False)]
    bufs = [(gg.DeviceArray(gpu, np.int32, G.n), gg.DeviceArray(gpu, np.int32, G.n)) for _ in range(2)]
    for i, s in enumerate(srcs):
        want = O.bfs(G.offsets, G.cols, s)[0]
        if i % 4 == 3:
            dg.config(mode=gg.GB_MODE_ASYNC)
            np.testing.assert_array_equal(dg.bfs(s)[0], want)
            dg.config(mode=gg.GB_MODE_GRAPH)
            continue
        lv, par = bufs[i % 2]
        dg.bfs_device(s, lv, par if i % 3 == 0 else None)
        np.testing.assert_array_equal(lv.numpy(), want, err_msg=f"replay {i} source {s}")
        if i % 3 == 0:
            assert O.check_parents(G.offsets, G.cols, s, want, par.numpy()) == 0


def test_isolated_and_degree_zero_sources(gpu):
    G = gg.Graph.rmat(12, seed=3)
    deg = np.diff(G.offsets)
    dg = gg.DeviceGraph(G, gpu)
    s = int(np.nonzero(deg == 0)[0][0])
    lv, par, st = dg.bfs(s, parents=True)
    assert lv[s] == 0 and (np.delete(lv, s) == -1).all()
    assert par[s] == s and st["n_reached"] == 1 and st["num_levels"] == 1 and st["m_reached"] == 0


def test_grid_many_levels(gpu):
    G = gg.Graph.grid(300, 200, 0.05, seed=4)
    dg = gg.DeviceGraph(G, gpu)
    for mode, dense in ((gg.GB_MODE_ASYNC, 0), (gg.GB_MODE_LEVEL_LOOP, 0), (gg.GB_MODE_ASYNC, 1)):
        dg.config(mode=mode, dense_edges=dense)
        for s in (0, 123 * 300 + 77):
            lv, _, st = dg.bfs(s)
            want, _, ost = O.bfs(G.offsets, G.cols, s)
            np.testing.assert_array_equal(lv, want)
            assert st["num_levels"] == ost["num_levels"] and st["num_levels"] > 300


def test_hub_spanning_many_tiles(gpu):
    # star with 100,000 leaves plus a tail: the hub's row covers ~100 tiles
    n = 100_002
    src = np.concatenate([np.zeros(100_000, np.int32), [100_000]]).astype(np.int32)
    dst = np.concatenate([np.arange(1, 100_001, dtype=np.int32), [100_001]]).astype(np.int32)
    G = gg.Graph.from_edges(n, np.concatenate([src, dst]), np.concatenate([dst, src]))
    dg = gg.DeviceGraph(G, gpu)
    for block, k, dense in [(256, 4, 0), (64, 16, 0), (1024, 2, 0), (256, 4, 1), (64, 16, 1)]:
        dg.config(block=block, edges_per_thread=k, dense_edges=dense)
        lv, par, _ = dg.bfs(5, parents=True)
        want, _, _ = O.bfs(G.offsets, G.cols, 5)
        np.testing.assert_array_equal(lv, want)
        assert lv.max() == 3  # leaf 5 -> hub 0 -> leaf 100000 -> tail 100001
        assert O.check_parents(G.offsets, G.cols, 5, lv, par) == 0


def test_tiny_graphs_and_errors(gpu):
    G = gg.Graph.from_edges(1, [], [])
    dg = gg.DeviceGraph(G, gpu)
    lv, _, st = dg.bfs(0)
    assert lv.tolist() == [0] and st["num_levels"] == 1
    with pytest.raises(gg.GBError) as e:
        dg.bfs(1)
    assert e.value.code == -5
    with pytest.raises(gg.GBError) as e:
        dg.bfs(-1)
    assert e.value.code == -5
    with pytest.raises(gg.GBError):
        dg.config(block=96, edges_per_thread=3)
    # self-loops and duplicates in a kept-order (reference layout) graph
    G = gg.Graph.from_edges(4, [0, 0, 0, 1, 1, 2], [0, 1, 1, 1, 2, 3], flags=gg.GB_KEEP_ORDER)
    lv, _, _ = gg.DeviceGraph(G, gpu).bfs(0)
    assert lv.tolist() == [0, 1, 2, 3]


def test_device_pointer_path_and_determinism(gpu):
    G = gg.Graph.rmat(15, seed=8)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    s = int(np.argmax(deg))
    d_lv = gg.DeviceArray(gpu, np.int32, G.n)
    d_par = gg.DeviceArray(gpu, np.int32, G.n)
    st = dg.bfs_device(s, d_lv, d_par)
    lv = d_lv.numpy()
    want, _, ost = O.bfs(G.offsets, G.cols, s)
    np.testing.assert_array_equal(lv, want)
    assert O.check_parents(G.offsets, G.cols, s, lv, d_par.numpy()) == 0
    assert st["n_reached"] == ost["n_reached"]
    for _ in range(5):
        np.testing.assert_array_equal(dg.bfs(s)[0], want)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 4095, 4096, 4097, 1_000_003])
def test_exclusive_scan(gpu, n):
    rng = np.random.default_rng(n)
    x = rng.integers(-2**31, 2**31 - 1, size=n, dtype=np.int64).astype(np.int32)
    if n == 0:
        d = gg.DeviceArray(gpu, np.int32, 1)
        gg._ffi.check(gg.lib().gb_exclusive_scan_i32(gpu, d.ptr, d.ptr, 0))
        return
    np.testing.assert_array_equal(gg.exclusive_scan(x, gpu), O.exclusive_scan(x))


def test_scan_known_answer(gpu):
    # 1D Partition/misc/test_prefix.cu:91-93
    assert gg.exclusive_scan(np.arange(16), gpu).tolist() == [0, 0, 1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 66, 78,
                                                               91, 105]


class _View:
    """A device pointer into another allocation (at an element offset)."""

    def __init__(self, base, offset):
        import ctypes
        self.ptr = ctypes.c_void_p(base.ptr.value + offset * base.dtype.itemsize)


def test_level_zero_in_the_clear_kernel(gpu):
    # Direct launches on symmetric graphs run level 0 inside the clear kernel
    # when the source has at most 128 CSR entries: sources of 127, 128 and
    # 129 entries, a row with every neighbour twice and a self-loop (rows in
    # file order), outputs at each 4-byte offset (16-byte stores or entry
    # stores), with and without parents, with the tail off, at level 1 and
    # learned; the replayed and level-loop modes for the same sources.
    rng = np.random.default_rng(91)
    base = gg.Graph.rmat(12, seed=91)
    n0 = base.n
    rows = np.repeat(np.arange(n0), np.diff(base.offsets))
    one_way = rows < base.cols
    src = [rows[one_way].astype(np.int64)]
    dst = [base.cols[one_way].astype(np.int64)]
    specials = []
    nxt = n0
    for d in (127, 128, 129, 3):
        nb = rng.choice(n0, d, replace=False)
        src.append(np.full(d, nxt))
        dst.append(nb)
        specials.append(nxt)
        nxt += 1
    dup = nxt
    nxt += 1
    nb = rng.choice(n0, 40, replace=False)
    src += [np.full(81, dup)]
    dst += [np.concatenate([nb, nb, [dup]])]
    specials.append(dup)
    G = gg.Graph.from_edges(nxt, np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32),
                            flags=gg.GB_SYMMETRIZE | gg.GB_KEEP_ORDER)
    assert G.is_symmetric
    deg = np.diff(G.offsets)
    assert [int(deg[v]) for v in specials] == [127, 128, 129, 3, 82]
    cand = np.nonzero(deg)[0]
    sources = specials + [int(x) for x in rng.choice(cand, 8, replace=False)]
    want = {s: O.bfs(G.offsets, G.cols, s)[0] for s in sources}
    dg = gg.DeviceGraph(G, gpu)
    lv_buf = gg.DeviceArray(gpu, np.int32, G.n + 8)
    par_buf = gg.DeviceArray(gpu, np.int32, G.n + 8)
    for tail in (gg.GB_TAIL_OFF, 1, gg.GB_TAIL_AUTO):
        dg.tail_config(tail)
        for parents in (False, True):
            for off in range(4):
                log(f"level 0 in the clear: tail {tail} parents {parents} offset {off}")
                for s in sources:
                    lv_buf.upload(np.full(lv_buf.count, 0x5A5A5A5A, dtype=np.int32))
                    dg.bfs_device(s, _View(lv_buf, off), _View(par_buf, off) if parents else None)
                    got = lv_buf.numpy()
                    np.testing.assert_array_equal(got[off:off + G.n], want[s], err_msg=f"source {s}")
                    assert (got[:off] == 0x5A5A5A5A).all() and (got[off + G.n:] == 0x5A5A5A5A).all()
                    if parents:
                        par = par_buf.numpy()[off:off + G.n]
                        assert O.check_parents(G.offsets, G.cols, s, want[s], par) == 0, s
    dg.tail_config(gg.GB_TAIL_AUTO)
    for mode in (gg.GB_MODE_GRAPH, gg.GB_MODE_LEVEL_LOOP):
        cfg = dg.get_config()
        cfg["mode"] = mode
        dg.config(**cfg)
        for s in specials:
            lv, _, _ = dg.bfs(s)
            np.testing.assert_array_equal(lv, want[s], err_msg=f"mode {mode} source {s}")
