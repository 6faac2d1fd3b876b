"""GPU parity of the search tail (bfs_tail_kernel): from a learned or set
level, one workgroup runs the rest of a single-source search on a symmetric
graph.  Levels, parents and statistics equal the oracle's (the restatement of
the reference's cpuBFS::cpuBFS, 1D Partition/src/cpuBFS.cpp:5-62) whatever
the tail's outcome: finished in the tail, declined at its first level (too
large for one workgroup: the per-level kernels go on from there), or handed
back at a later level (the deep-mode kernel goes on from there)."""
import numpy as np
import pytest

import gpu_graphs_amd as gg
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def check_search(G, dg, s):
    lv, par, st = dg.bfs(int(s), parents=True)
    want, _, ost = O.bfs(G.offsets, G.cols, int(s))
    np.testing.assert_array_equal(lv, want, err_msg=f"source {s}")
    assert O.check_parents(G.offsets, G.cols, int(s), lv, par) == 0, s
    assert (st["n_reached"], st["m_reached"], st["num_levels"]) == (
        ost["n_reached"], ost["m_reached"], ost["num_levels"]), s
    return dg.get_tail_config()


def symmetric(n, pairs):
    a = np.array(pairs, dtype=np.int32).reshape(-1, 2)
    src = np.concatenate([a[:, 0], a[:, 1]])
    dst = np.concatenate([a[:, 1], a[:, 0]])
    return gg.Graph.from_edges(n, src, dst)


def broom():
    """Path 0..19; vertex 19 - hub 20 with 3000 leaves (21..3020, level 21
    from vertex 0: more entries than one workgroup keeps); leaf 3020 starts
    a second path of 40 vertices (levels 22..61)."""
    pairs = [(i, i + 1) for i in range(19)] + [(19, 20)]
    pairs += [(20, 21 + j) for j in range(3000)]
    pairs += [(3020, 3021)] + [(3021 + i, 3022 + i) for i in range(39)]
    return symmetric(3061, pairs)


def two_hubs():
    """Path 0..9; vertex 9 - hubs 10 and 11 with 20000 leaves each: level 10
    holds 2 vertices but more CSR entries than the tail expands."""
    pairs = [(i, i + 1) for i in range(9)] + [(9, 10), (9, 11), (10, 11)]
    pairs += [(10, 12 + j) for j in range(20000)] + [(11, 20012 + j) for j in range(20000)]
    return symmetric(40012, pairs)


@pytest.mark.parametrize("mode", [gg.GB_MODE_GRAPH, gg.GB_MODE_ASYNC])
def test_tail_outcomes_on_a_broom(gpu, mode):
    G = broom()
    dg = gg.DeviceGraph(G, gpu)
    dg.config(mode=mode)
    dg.deep_config(gg.GB_DEEP_OFF)
    # (tail level, source, expected outcome)
    cases = [(5, 0, gg.GB_TAIL_HANDED_BACK),   # level 21 is too large: the deep kernel goes on
             (0, 0, gg.GB_TAIL_HANDED_BACK),
             (21, 0, gg.GB_TAIL_DECLINED),     # its first level is too large
             (22, 0, gg.GB_TAIL_FINISHED),     # the second path, 40 levels in one workgroup
             (0, 3060, gg.GB_TAIL_HANDED_BACK),   # the other way: the hub at level 41, its leaves at 42
             (30, 3060, gg.GB_TAIL_HANDED_BACK),
             (1, 25, gg.GB_TAIL_HANDED_BACK),  # from a leaf: level 1 is the hub, level 2 its leaves
             (3, 25, gg.GB_TAIL_FINISHED)]
    for level, s, outcome in cases:
        dg.tail_config(level)
        t = check_search(G, dg, s)
        log(f"broom mode={mode} tail at {level} from {s}: {t}")
        assert t == {"level": level, "used": level, "outcome": outcome}, (level, s)
    dg.tail_config(gg.GB_TAIL_OFF)
    assert check_search(G, dg, 0)["used"] == -1


def test_tail_hands_back_a_level_with_too_many_entries(gpu):
    G = two_hubs()
    dg = gg.DeviceGraph(G, gpu)
    dg.deep_config(gg.GB_DEEP_OFF)
    for level, s, outcome in [(2, 0, gg.GB_TAIL_HANDED_BACK), (0, 0, gg.GB_TAIL_HANDED_BACK),
                              (11, 0, gg.GB_TAIL_DECLINED),
                              (12, 0, gg.GB_TAIL_FINISHED),  # level 12 is empty: the tail ends the search
                              (13, 0, 0)]:                   # the search ended before the tail's level
        dg.tail_config(level)
        t = check_search(G, dg, s)
        log(f"two hubs: tail at {level}: {t}")
        assert t["outcome"] == outcome, (level, t)


def test_long_path_in_one_workgroup(gpu):
    # 3000 levels of one vertex each, all in the tail.
    n = 3000
    G = symmetric(n, [(i, i + 1) for i in range(n - 1)])
    dg = gg.DeviceGraph(G, gpu)
    dg.deep_config(gg.GB_DEEP_OFF)
    dg.tail_config(0)
    for s in (0, 1500, n - 1):
        assert check_search(G, dg, s)["outcome"] == gg.GB_TAIL_FINISHED


@pytest.mark.parametrize("scale,seed", [(14, 101), (17, 102)])
def test_learned_tail_on_rmat(gpu, scale, seed):
    # The learned tail engages after a few searches and every result stays
    # exact, whatever level each source's small levels start at.
    G = gg.Graph.rmat(scale, seed=seed)
    dg = gg.DeviceGraph(G, gpu)
    deg = np.diff(G.offsets)
    nz = np.nonzero(deg)[0]
    rng = np.random.default_rng(seed)
    sources = list(rng.choice(nz, 24, replace=False)) + [int(np.argmax(deg))] + list(np.nonzero(deg == 1)[0][:3])
    used = []
    for s in sources:
        t = check_search(G, dg, s)
        used.append(t["used"])
    log(f"rmat{scale}: tail levels used {used}")
    assert max(used[4:]) >= 0  # the tail ran
    for level in (1, 2, 3):  # set levels: declined or handed back where the frontier is large
        dg.tail_config(level)
        for s in sources[:6]:
            check_search(G, dg, s)


def test_tail_not_used_where_it_does_not_apply(gpu):
    # Directed graphs (two launches per level), deep mode and the host level
    # loop never start a tail.
    G = gg.Graph.rmat(13, seed=103)
    rows = np.repeat(np.arange(G.n, dtype=np.int32), np.diff(G.offsets))
    up = rows < G.cols
    D = gg.Graph.from_edges(G.n, rows[up], G.cols[up])
    for graph, setup in [(D, lambda d: None), (G, lambda d: d.deep_config(gg.GB_DEEP_ON)),
                         (G, lambda d: d.config(mode=gg.GB_MODE_LEVEL_LOOP))]:
        dg = gg.DeviceGraph(graph, gpu)
        setup(dg)
        dg.tail_config(0)
        for s in (0, 5, 77):
            assert check_search(graph, dg, s)["used"] == -1
