"""GPU parity of the batched multi-source BFS (gb_bfs_batch): every source's
level row must be bit-exact with the single-source path and the CPU oracle,
with and without hipGraph replay, for batches of 1..64 sources (repeats
allowed), on graphs with sinks, deep graphs and hubs."""
import numpy as np
import pytest

import gpu_graphs_amd as gg
from conftest import load_fixture
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(256, 1), (256, 2), (256, 4), (512, 2), (1024, 1), (64, 16)]


def log(msg):
    print(msg, flush=True)


def check_batch(G, dg, sources):
    levels, st = dg.bfs_batch(sources)
    want = O.bfs_multi(G.offsets, G.cols, sources)
    np.testing.assert_array_equal(levels, want)
    deg = np.diff(G.offsets)
    for i, s in enumerate(sources):
        reached = want[i] >= 0
        assert st["n_reached"][i] == int(reached.sum())
        assert st["m_reached"][i] == int(deg[reached].sum())
    assert st["num_levels"] == int(want.max()) + 1
    return levels


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("block,k", SHAPES)
def test_rmat_batch_shapes(gpu, block, k, use_graph):
    G = gg.Graph.rmat(14, seed=51)
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_config(block=block, edges_per_thread=k, use_graph=use_graph)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(block + k)
    sources = rng.choice(np.nonzero(deg)[0], 64, replace=False).astype(np.int32)
    log(f"rmat14 batch block={block} k={k} graph={use_graph}")
    check_batch(G, dg, sources)


@pytest.mark.parametrize("nsrc", [1, 2, 17, 63, 64])
def test_batch_sizes_with_repeats_and_isolated(gpu, nsrc):
    G = gg.Graph.rmat(13, seed=7)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(nsrc)
    pool = np.concatenate([np.nonzero(deg)[0][:40], np.nonzero(deg == 0)[0][:3]])
    sources = rng.choice(pool, nsrc, replace=True).astype(np.int32)
    check_batch(G, dg, sources)


def test_batch_equals_single_source(gpu):
    G = gg.Graph.rmat(16, seed=3)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    sources = np.random.default_rng(0).choice(np.nonzero(deg)[0], 64, replace=False).astype(np.int32)
    levels, _ = dg.bfs_batch(sources)
    for i in range(0, 64, 7):
        np.testing.assert_array_equal(levels[i], dg.bfs(int(sources[i]))[0])
    np.testing.assert_array_equal(levels, O.bfs_multi(G.offsets, G.cols, sources))


def test_directed_sinks_on_scanned_levels(gpu):
    # Vertices without out-edges (sinks) reached by levels whose update scans
    # all vertices: on a directed graph they are reachable, unlike the
    # edgeless vertices of a symmetric graph, and the scan must visit them.
    n = 5000
    rng = np.random.default_rng(12)
    src = np.concatenate([np.zeros(2000, np.int64), rng.integers(0, 100, 8000)]).astype(np.int32)
    dst = np.concatenate([np.arange(1, 2001), rng.integers(100, n, 8000)]).astype(np.int32)
    G = gg.Graph.from_edges(n, src, dst)
    assert not G.is_symmetric
    deg = np.diff(G.offsets)
    assert (deg == 0).sum() > 1000
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_config(sparse_edges=1)  # every level with more than one edge is dense: its update scans
    for sources in (np.array([0, 1, 2, 50, 99], np.int32), np.arange(64, dtype=np.int32)):
        log(f"directed sinks, {len(sources)} sources")
        levels, _ = dg.bfs_batch(sources)
        np.testing.assert_array_equal(levels, O.bfs_multi(G.offsets, G.cols, sources))


@pytest.mark.parametrize("name", ["wiki-Vote", "test2", "tree", "test_isolated_node"])
def test_fixture_batches(gpu, name, golden_levels):
    g = load_fixture(name)
    dg = g.device_graph(gpu)
    sources = golden_levels[f"{name}__sources"]
    want = golden_levels[f"{name}__levels"]
    for lo in range(0, len(sources), 64):
        lv, _ = dg.bfs_batch(sources[lo:lo + 64])
        np.testing.assert_array_equal(lv, want[lo:lo + 64])


def test_deep_grid_beyond_depth_hint(gpu):
    # hundreds of levels: the captured graph covers the depth hint, the rest
    # runs as continuation launches; the second batch reuses the learned depth.
    G = gg.Graph.grid(120, 90, 0.05, seed=9)
    dg = gg.DeviceGraph(G, gpu)
    sources = np.array([0, 5, 120 * 45 + 60, G.n - 1] * 4, dtype=np.int32)
    for _ in range(2):
        check_batch(G, dg, sources)


# sparse_edges: 1 = every level scans all vertices, 2**32-1 = every level
# hands a touched list to the next, 0 = the default threshold.
@pytest.mark.parametrize("sparse", [1, 2**32 - 1, 0])
@pytest.mark.parametrize("graph", ["rmat", "grid"])
def test_level_modes(gpu, graph, sparse):
    if graph == "rmat":
        G = gg.Graph.rmat(14, seed=61)
    else:
        G = gg.Graph.grid(150, 70, 0.05, seed=62)
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_config(sparse_edges=sparse)
    assert dg.get_batch_config()["sparse_edges"] == (sparse or max(1 << 16, G.n // 16))
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(sparse % 1000)
    sources = rng.choice(np.nonzero(deg)[0], 40, replace=False).astype(np.int32)
    log(f"{graph} n={G.n} sparse_edges={sparse}")
    for _ in range(2):  # second run replays the captured sequence
        check_batch(G, dg, sources)


def test_one_capture_serves_any_k_and_buffer(gpu):
    # Sources, k and the output buffer live in device memory: batches of
    # different sizes into different buffers replay one captured sequence.
    G = gg.Graph.rmat(13, seed=63)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(63)
    cand = np.nonzero(deg)[0]
    bufs = [gg.DeviceArray(gpu, np.int32, 64 * G.n) for _ in range(2)]
    for rep, k in enumerate([64, 5, 1, 33, 64]):
        sources = rng.choice(cand, k, replace=False).astype(np.int32)
        d = bufs[rep % 2]
        dg.bfs_batch_device(sources, d)
        got = d.numpy_slice(0, k * G.n).reshape(k, G.n)
        np.testing.assert_array_equal(got, O.bfs_multi(G.offsets, G.cols, sources), err_msg=f"k={k}")


def test_graph_replays_match_eager(gpu):
    G = gg.Graph.rmat(15, seed=12)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(4)
    batches = [rng.choice(np.nonzero(deg)[0], 64, replace=False).astype(np.int32) for _ in range(3)]
    dg.batch_config(use_graph=False)
    eager = [dg.bfs_batch(b, stats=False)[0] for b in batches]
    dg.batch_config(use_graph=True)
    d_lv = gg.DeviceArray(gpu, np.int32, 64 * G.n)
    for rep in range(4):
        for b, want in zip(batches, eager):
            dg.bfs_batch_device(b, d_lv)
            np.testing.assert_array_equal(d_lv.numpy().reshape(64, G.n), want)
            # eager single-source work between replays
            np.testing.assert_array_equal(dg.bfs(int(b[rep]))[0], want[rep])


def test_batch_errors(gpu):
    G = gg.Graph.rmat(10, seed=1)
    dg = gg.DeviceGraph(G, gpu)
    with pytest.raises(gg.GBError) as e:
        dg.bfs_batch(np.zeros(65, dtype=np.int32))
    assert e.value.code == -1
    with pytest.raises(gg.GBError) as e:
        dg.bfs_batch(np.array([0, G.n], dtype=np.int32))
    assert e.value.code == -5
    with pytest.raises(gg.GBError):
        dg.batch_config(block=100, edges_per_thread=1)


@pytest.mark.parametrize("alpha", [1, 2, 8, 1000])
@pytest.mark.parametrize("enable", [True, False])
def test_bottom_up_batches(gpu, enable, alpha):
    # Bottom-up levels pull frontier masks from neighbours; every row must
    # equal the oracle's whatever the switch point.
    G = gg.Graph.rmat(16, seed=81)
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_direction_config(enable, alpha)
    assert dg.get_batch_direction_config() == {"enable": enable, "alpha": alpha, "beta": 0}
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(alpha)
    for k in (64, 7):
        sources = rng.choice(np.nonzero(deg)[0], k, replace=False).astype(np.int32)
        log(f"rmat16 bottom-up={enable} alpha={alpha} k={k}")
        check_batch(G, dg, sources)
        if not enable:
            assert dg.batch_last_bottom_up() == 0
        elif alpha >= 2 and k == 64:
            assert dg.batch_last_bottom_up() >= 1


def test_bottom_up_heavy_rows_and_small_components(gpu):
    # Rows of thousands of entries are scanned in tiles spread over all
    # workgroups (several tiles per row, one row spanning the tile boundary
    # of the next), and sources in small components finish early while the
    # others still search.
    rng = np.random.default_rng(85)
    base = gg.Graph.rmat(14, seed=85)
    n0 = base.n
    src = [np.repeat(np.arange(n0), np.diff(base.offsets)).astype(np.int64)]
    dst = [base.cols.astype(np.int64)]
    for h, d in ((0, 5000), (1, 2048), (2, 1024), (3, 1023), (4, 3333)):
        nb = rng.choice(n0, d, replace=False)
        src += [np.full(d, h), nb]
        dst += [nb, np.full(d, h)]
    # ten 3-vertex paths, each its own component
    for c in range(10):
        a, b, c3 = n0 + 3 * c, n0 + 3 * c + 1, n0 + 3 * c + 2
        src += [np.array([a, b, b, c3])]
        dst += [np.array([b, a, c3, b])]
    G = gg.Graph.from_edges(n0 + 30, np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32))
    assert G.is_symmetric and int(np.diff(G.offsets).max()) >= 5000
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_direction_config(True, 8)
    deg = np.diff(G.offsets)
    giant = rng.choice(np.nonzero(deg[:n0])[0], 54, replace=False)
    sources = np.concatenate([giant, n0 + 3 * np.arange(10)]).astype(np.int32)
    log("heavy rows + small components")
    check_batch(G, dg, sources)
    assert dg.batch_last_bottom_up() >= 1


def test_bottom_up_batch_on_mixed_graph_and_directed(gpu):
    # R-MAT core plus a grid tail joined by one edge: bottom-up big levels,
    # then thin levels; and a directed graph, where bottom-up must stay off.
    core = gg.Graph.rmat(13, seed=82)
    grid = gg.Graph.grid(120, 40, 0.05, seed=83)
    nc = core.n
    cs = np.repeat(np.arange(nc), np.diff(core.offsets)).astype(np.int32)
    gs = (np.repeat(np.arange(grid.n), np.diff(grid.offsets)) + nc).astype(np.int32)
    hub = int(np.argmax(np.diff(core.offsets)))
    G = gg.Graph.from_edges(nc + grid.n, np.concatenate([cs, gs, [hub, nc]]),
                            np.concatenate([core.cols, grid.cols + nc, [nc, hub]]))
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    sources = np.random.default_rng(82).choice(np.nonzero(deg)[0], 48, replace=False).astype(np.int32)
    check_batch(G, dg, sources)
    assert dg.batch_last_bottom_up() >= 1
    rng = np.random.default_rng(84)
    n = 5000
    D = gg.Graph.from_edges(n, rng.integers(0, n, 40000), rng.integers(0, n, 40000))
    dd = gg.DeviceGraph(D, gpu)
    check_batch(D, dd, np.arange(0, 64, dtype=np.int32) * 7)
    assert dd.batch_last_bottom_up() == 0


def test_bottom_up_window_carried_to_other_batches(gpu):
    # A runner remembers the levels its last batch's direction rule took
    # bottom-up; the next batch (same thresholds) runs those levels
    # bottom-up after a scanning update whatever its own frontiers say, and
    # their update builds no queue.  Batches whose own choice differs (other
    # sources, fewer of them, sources in the thin tail) must still equal the
    # oracle, and so must a batch after the thresholds change.
    core = gg.Graph.rmat(14, seed=87)
    grid = gg.Graph.grid(150, 30, 0.05, seed=88)
    nc = core.n
    cs = np.repeat(np.arange(nc), np.diff(core.offsets)).astype(np.int32)
    gs = (np.repeat(np.arange(grid.n), np.diff(grid.offsets)) + nc).astype(np.int32)
    hub = int(np.argmax(np.diff(core.offsets)))
    G = gg.Graph.from_edges(nc + grid.n, np.concatenate([cs, gs, [hub, nc]]),
                            np.concatenate([core.cols, grid.cols + nc, [nc, hub]]))
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_direction_config(True, 4, 0)
    deg = np.diff(G.offsets[: nc + 1])
    rng = np.random.default_rng(87)
    by_degree = np.argsort(-deg, kind="stable")
    hubs = by_degree[:64].astype(np.int32)
    low = rng.choice(np.nonzero(deg == 1)[0], 48, replace=False).astype(np.int32)
    tail = (nc + rng.choice(grid.n, 16, replace=False)).astype(np.int32)
    for name, sources in (("hubs", hubs), ("low-degree + tail", np.concatenate([low, tail])),
                          ("three", hubs[:3]), ("tail", tail), ("hubs again", hubs)):
        log(f"window batch: {name}")
        check_batch(G, dg, sources)
    assert dg.batch_last_bottom_up() >= 1
    dg.batch_direction_config(True, 1, 0)
    log("window batch: alpha changed")
    check_batch(G, dg, np.concatenate([low, tail]))


def test_level_lines_of_earlier_batches_are_not_read(gpu):
    # The per-vertex level lines are not cleared between batches: the first
    # level that reaches a vertex rewrites its whole line, and vertices no
    # source reaches must come out -1 whatever an earlier batch left there.
    base = gg.Graph.rmat(12, seed=86)
    n0 = base.n
    src = [np.repeat(np.arange(n0), np.diff(base.offsets)).astype(np.int64)]
    dst = [base.cols.astype(np.int64)]
    for c in range(10):  # ten 3-vertex paths, each its own component
        a, b, c3 = n0 + 3 * c, n0 + 3 * c + 1, n0 + 3 * c + 2
        src += [np.array([a, b, b, c3])]
        dst += [np.array([b, a, c3, b])]
    G = gg.Graph.from_edges(n0 + 30, np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32))
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    rng = np.random.default_rng(86)
    giant = rng.choice(np.nonzero(deg[:n0])[0], 64, replace=False).astype(np.int32)
    paths = (n0 + 3 * np.arange(10)).astype(np.int32)
    batches = [
        np.concatenate([giant[:54], paths]),        # every component reached
        np.concatenate([paths[::-1], giant[10:]]),  # the paths' lines now under other sources
        giant[:20],                                 # the paths unreached: their old lines stay behind
        paths[[4, 1, 7]],                           # only three paths: the giant component unreached
        giant[30:64],
    ]
    for i, sources in enumerate(batches):
        log(f"stale lines, batch {i} ({len(sources)} sources)")
        check_batch(G, dg, sources)
    # Directed: every vertex is scanned on dense levels (sinks included), and
    # which vertices a batch reaches depends on where its sources sit.
    rng = np.random.default_rng(87)
    n = 6000
    src = np.concatenate([rng.integers(0, 3000, 30000), rng.integers(3000, n, 4000)])
    dst = np.concatenate([rng.integers(0, n, 30000), rng.integers(3000, n, 4000)])
    D = gg.Graph.from_edges(n, src.astype(np.int32), dst.astype(np.int32))
    assert not D.is_symmetric
    dd = gg.DeviceGraph(D, gpu)
    dd.batch_config(sparse_edges=1)  # dense levels: the updates scan
    for i, sources in enumerate((np.arange(0, 6000, 94, dtype=np.int32)[:64],  # both halves
                                 np.arange(3000, 3064, dtype=np.int32),       # the half that cannot reach [0, 3000)
                                 np.arange(0, 10, dtype=np.int32))):
        log(f"stale lines, directed batch {i}")
        check_batch(D, dd, sources)


@pytest.mark.parametrize("use_graph", [True, False])
def test_edgeless_sources_across_batches(gpu, use_graph):
    # After the first batch only vertices with edges are cleared; an edgeless
    # source is cleared by the next batch's seed.  Its stale state would show
    # as a wrong row or count under the source that takes its bit next.
    G = gg.Graph.rmat(13, seed=7)
    assert G.is_symmetric
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_config(use_graph=use_graph)
    deg = np.diff(G.offsets)
    iso = np.nonzero(deg == 0)[0][:4].astype(np.int32)
    normal = np.nonzero(deg)[0][:60].astype(np.int32)
    assert iso.size == 4
    batches = [
        np.concatenate([normal[:3], iso[:1], normal[3:20]]),          # iso[0] at bit 3
        np.concatenate([normal[20:30], iso[:1], normal[30:40]]),      # bit 3 now a normal source, iso[0] at bit 10
        normal[:64],                                                  # no edgeless source
        np.array([iso[1], iso[1], iso[2], normal[5]], dtype=np.int32),  # repeats
        np.concatenate([iso, normal[:8]]),
        normal[10:50],
    ]
    for i, sources in enumerate(batches):
        log(f"edgeless sources, batch {i} ({len(sources)} sources)")
        check_batch(G, dg, sources)


@pytest.mark.parametrize("use_graph", [True, False])
def test_fused_deep_batches(gpu, use_graph):
    # Fused levels (one launch per level, top-down only), forced on: a deep
    # grid (levels far past the 254 kept in bytes), an R-MAT graph (wide
    # levels, hub rows pushed by whole waves), a directed path, and graphs
    # with edgeless and repeated sources, batch after batch on one handle.
    cases = []
    Gg = gg.Graph.grid(150, 120, 0.05, seed=95)
    cases.append(("grid", Gg, [np.array([0, 7, Gg.n - 1, 9000], dtype=np.int32),
                               np.arange(0, Gg.n, Gg.n // 64, dtype=np.int32)[:64],
                               np.array([5, 5, 17000, 5], dtype=np.int32)]))
    R = gg.Graph.rmat(13, seed=7)
    deg = np.diff(R.offsets)
    iso = np.nonzero(deg == 0)[0][:2].astype(np.int32)
    hub = np.array([np.argmax(deg)], dtype=np.int32)
    nz = np.nonzero(deg)[0].astype(np.int32)
    cases.append(("rmat", R, [np.concatenate([nz[:30], iso, hub]), nz[30:94], np.concatenate([iso, iso, nz[:3]])]))
    n = 3000
    src = np.arange(n - 1, dtype=np.int32)
    D = gg.Graph.from_edges(n, src, src + 1)
    cases.append(("directed path", D, [np.array([0, 1500, n - 1], dtype=np.int32), np.arange(0, n, 50, dtype=np.int32)]))
    for name, G, batches in cases:
        dg = gg.DeviceGraph(G, gpu)
        dg.batch_config(use_graph=use_graph)
        dg.batch_deep_config(gg.GB_DEEP_ON)
        for i, sources in enumerate(batches):
            log(f"fused batches: {name} batch {i} ({len(sources)} sources)")
            check_batch(G, dg, sources)
            assert dg.get_batch_deep_config() == {"mode": gg.GB_DEEP_ON, "used": True}
        dg.batch_deep_config(gg.GB_DEEP_OFF)  # back to the two-launch levels on the same workspace
        check_batch(G, dg, batches[0])
        assert not dg.get_batch_deep_config()["used"]


def test_fused_batches_learned_on_deep_graphs_only(gpu):
    G = gg.Graph.grid(200, 150, 0.05, seed=96)
    dg = gg.DeviceGraph(G, gpu)
    used = []
    for i in range(3):
        sources = np.arange(i, G.n, G.n // 16, dtype=np.int32)[:16]
        check_batch(G, dg, sources)
        used.append(dg.get_batch_deep_config()["used"])
    assert used == [False, True, True], used
    R = gg.Graph.rmat(14, seed=97)
    dr = gg.DeviceGraph(R, gpu)
    nz = np.nonzero(np.diff(R.offsets))[0].astype(np.int32)
    for i in range(3):
        check_batch(R, dr, nz[i * 20:(i + 1) * 20])
        assert not dr.get_batch_deep_config()["used"]


def test_levels_past_the_byte_range(gpu):
    # A 700-vertex path: levels go past 254, where the batch stores levels
    # into the rows directly instead of the byte lines.
    n = 700
    src = np.arange(n - 1, dtype=np.int32)
    G = gg.Graph.from_edges(n, np.concatenate([src, src + 1]), np.concatenate([src + 1, src]))
    dg = gg.DeviceGraph(G, gpu)
    sources = np.array([0, n - 1, 350, 253, 254, 255, 0], dtype=np.int32)
    levels = check_batch(G, dg, sources)
    assert levels[0, n - 1] == n - 1 and levels[2, 0] == 350


@pytest.mark.parametrize("use_graph", [True, False])
def test_shallow_lines_then_a_deeper_batch(gpu, use_graph):
    # After shallow batches a runner keeps levels in 4-bit lines (levels up
    # to 13; deeper ones go to the output rows directly).  A batch that turns
    # out deeper than that (sources on a separate 40-vertex path beside an
    # R-MAT core) stays exact, and so do the batches after it, which go back
    # to byte lines and then to 4-bit lines again.
    core = gg.Graph.rmat(13, seed=98)
    nc, plen = core.n, 40
    cs = np.repeat(np.arange(nc), np.diff(core.offsets))
    path = np.arange(nc, nc + plen)
    src = np.concatenate([cs, path[:-1], path[1:]]).astype(np.int32)
    dst = np.concatenate([core.cols, path[1:], path[:-1]]).astype(np.int32)
    G = gg.Graph.from_edges(nc + plen, src, dst)
    dg = gg.DeviceGraph(G, gpu)
    dg.batch_config(use_graph=use_graph)
    nz = np.nonzero(np.diff(G.offsets)[:nc])[0].astype(np.int32)
    shallow = [nz[i * 40:(i + 1) * 40] for i in range(3)]
    deep = np.concatenate([nz[200:230], path[-10:].astype(np.int32), [nc + plen - 1, nc + plen - 1]])
    for i, sources in enumerate(shallow + [deep, shallow[0], shallow[1], deep[::-1].copy(), shallow[2]]):
        log(f"lines batch {i}: {len(sources)} sources")
        levels = check_batch(G, dg, sources)
        assert (levels.max() >= 30) == (i in (3, 6)), i


class _View:
    """A device pointer into another allocation (at an element offset)."""

    def __init__(self, base, offset):
        import ctypes
        self.ptr = ctypes.c_void_p(base.ptr.value + offset * base.dtype.itemsize)


def test_rows_at_any_alignment(gpu):
    # Output rows that start 16-byte aligned or not (a device buffer at an
    # offset of 0..3 entries), with levels kept in the lines and levels stored
    # into the rows directly (a 700-vertex path, levels past 254); n is a
    # multiple of 4 in both, the entries around the rows stay untouched.
    n = 700
    p = np.arange(n - 1, dtype=np.int32)
    path = gg.Graph.from_edges(n, np.concatenate([p, p + 1]), np.concatenate([p + 1, p]))
    cases = [(gg.Graph.rmat(12, seed=64), None), (path, np.array([0, n - 1, 350, 3], dtype=np.int32))]
    for G, sources in cases:
        assert G.n % 4 == 0
        if sources is None:
            cand = np.nonzero(np.diff(G.offsets))[0]
            sources = np.random.default_rng(64).choice(cand, 64, replace=False).astype(np.int32)
        k = sources.size
        want = O.bfs_multi(G.offsets, G.cols, sources)
        dg = gg.DeviceGraph(G, gpu)
        base = gg.DeviceArray(gpu, np.int32, k * G.n + 8)
        for off in range(4):
            log(f"alignment: n={G.n} k={k} offset {off}")
            base.upload(np.full(base.count, 0x5A5A5A5A, dtype=np.int32))
            dg.bfs_batch_device(sources, _View(base, off))
            got = base.numpy()
            np.testing.assert_array_equal(got[off:off + k * G.n].reshape(k, G.n), want, err_msg=f"offset {off}")
            assert (got[:off] == 0x5A5A5A5A).all() and (got[off + k * G.n:] == 0x5A5A5A5A).all()
