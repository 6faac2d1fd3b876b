"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against at least one of: the reference's own
known answers (the scan print test, the fixture graphs' closed-form levels),
the reference's own code compiled from its sources (oracle/_ref, when
built), an independent library (scipy BFS goldens), closed forms at size,
brute force on tiny inputs, and invariants chosen so that a dropped term or
a shifted index fails.
"""
import itertools

import numpy as np
import pytest

from conftest import FIXTURES, fixture_path
from oracle import oracle as O

# Closed-form levels of the reference fixtures (derived from the files' edge
# lists, independent of any BFS code): name -> (one_based, source, levels).
CLOSED_FORMS = [
    ("tree", False, 1, [-1, 0, 1, 1, 2, 2, 2, 2]),
    ("test4", False, 1, [-1, 0, 1, 2, 3, 4, -1, -1]),
    ("test3", True, 0, [0, 1, 1, 1, 1]),
    ("test3", True, 4, [-1, -1, -1, -1, 0]),
    ("test_isolated_node", True, 1, [-1, 0, 1, 2]),
    ("test_isolated_node", True, 0, [0, -1, -1, -1]),
    ("mig_tree", True, 0, [0, 1, 1, 2, 2, 2, 2]),
]


def csr_from_pairs(n, pairs):
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    order = np.argsort(pairs[:, 0], kind="stable")
    p = pairs[order]
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, p[:, 0] + 1, 1)
    return np.cumsum(off), p[:, 1].astype(np.int32)


def undirected(n, pairs):
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    return csr_from_pairs(n, np.concatenate([pairs, pairs[:, ::-1]]))


@pytest.mark.parametrize("name,one_based,source,expected", CLOSED_FORMS)
def test_closed_form_fixtures(name, one_based, source, expected):
    off, cols = O.load_edgelist(fixture_path(name), one_based=one_based)
    lv, _, st = O.bfs(off, cols, source)
    assert lv.tolist() == expected
    assert st["num_levels"] == max(expected) + 1
    assert st["n_reached"] == sum(1 for x in expected if x >= 0)


def test_threshold_fixtures_closed_form():
    # test.txt: 1 -> 2..1029, 2 -> 1030..2058 (1-based): from 0-based source 0,
    # 1028 vertices at level 1 and 1029 at level 2 (crosses 1024).
    for name in ("test", "test2"):
        off, cols = O.load_edgelist(fixture_path(name), one_based=True)
        lv, _, _ = O.bfs(off, cols, 0)
        assert lv[0] == 0
        assert np.count_nonzero(lv == 1) == 1028 and np.count_nonzero(lv == 2) == 1029
        assert (lv >= 0).all()


@pytest.mark.parametrize("name", list(FIXTURES))
def test_oracle_matches_scipy_goldens(name, golden_levels):
    one_based, auto_n = FIXTURES[name]
    if auto_n:
        pytest.skip("the oracle loader keeps the header n; covered through the engine loader")
    off, cols = O.load_edgelist(fixture_path(name), one_based=one_based)
    sources = golden_levels[f"{name}__sources"]
    want = golden_levels[f"{name}__levels"]
    got = O.bfs_multi(off, cols, sources)
    np.testing.assert_array_equal(got, want)


def test_oracle_wiki_vote_matches_scipy(golden_levels):
    # ids exceed the header n (8297 > 7115): load with n = max id + 1 via the engine.
    import gpu_graphs_amd as gg
    g = gg.csr(fixture_path("wiki-Vote"), one_based=True, auto_n=True)
    got = O.bfs_multi(g.row_offset, g.col_idx, golden_levels["wiki-Vote__sources"])
    np.testing.assert_array_equal(got, golden_levels["wiki-Vote__levels"])


def test_er_golden_levels():
    import os

    from conftest import GOLDEN
    off, cols = O.load_edgelist(os.path.join(GOLDEN, "er_10200.txt"))
    want = np.load(os.path.join(GOLDEN, "er_10200_levels.npy"))
    lv, _, st = O.bfs(off, cols, 0)
    np.testing.assert_array_equal(lv, want)
    assert (lv[10000:] == -1).all()  # the 200 isolated ids


def test_dequeue_model_equals_discover_model():
    # cpuBFS::testDistance (cpuBFS.cpp:108-119): linear_bfs's visit-on-dequeue
    # labelling must give the same levels as discover-time labelling.
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(2, 200))
        m = int(rng.integers(0, 4 * n))
        pairs = rng.integers(0, n, size=(m, 2))
        off, cols = csr_from_pairs(n, pairs)
        s = int(rng.integers(0, n))
        np.testing.assert_array_equal(O.bfs_dequeue(off, cols, s), O.bfs(off, cols, s)[0])


def floyd_levels(n, pairs, s):
    INF = 10**9
    d = np.full((n, n), INF, dtype=np.int64)
    np.fill_diagonal(d, 0)
    for u, v in pairs:
        if u != v:
            d[u, v] = min(d[u, v], 1)
    for k in range(n):
        d = np.minimum(d, d[:, k:k + 1] + d[k:k + 1, :])
    row = d[s]
    return np.where(row >= INF, -1, row).astype(np.int32)


def test_brute_force_tiny_graphs():
    rng = np.random.default_rng(11)
    for trial in range(60):
        n = int(rng.integers(1, 40))
        m = int(rng.integers(0, 3 * n + 1))
        pairs = rng.integers(0, n, size=(m, 2))
        off, cols = csr_from_pairs(n, pairs)
        for s in {0, int(rng.integers(0, n))}:
            np.testing.assert_array_equal(O.bfs(off, cols, s)[0], floyd_levels(n, pairs, s))


def test_closed_forms_at_size():
    # path: |i - s|
    n = 5000
    off, cols = undirected(n, [(i, i + 1) for i in range(n - 1)])
    for s in (0, 1234, n - 1):
        np.testing.assert_array_equal(O.bfs(off, cols, s)[0], np.abs(np.arange(n) - s))
    # full W x H grid: Manhattan distance
    W, H = 70, 50
    ids = np.arange(W * H).reshape(H, W)
    e = np.concatenate([np.stack([ids[:, :-1].ravel(), ids[:, 1:].ravel()], 1),
                        np.stack([ids[:-1, :].ravel(), ids[1:, :].ravel()], 1)])
    off, cols = undirected(W * H, e)
    sy, sx = 17, 33
    want = (np.abs(np.arange(H)[:, None] - sy) + np.abs(np.arange(W)[None, :] - sx)).ravel()
    np.testing.assert_array_equal(O.bfs(off, cols, int(ids[sy, sx]))[0], want)
    # star, complete graph, two components
    off, cols = undirected(101, [(0, i) for i in range(1, 101)])
    lv = O.bfs(off, cols, 7)[0]
    assert lv[7] == 0 and lv[0] == 1 and (np.delete(lv, [0, 7]) == 2).all()
    k = 30
    off, cols = undirected(k, list(itertools.combinations(range(k), 2)))
    lv = O.bfs(off, cols, 3)[0]
    assert lv[3] == 0 and (np.delete(lv, 3) == 1).all()
    off, cols = undirected(10, [(0, 1), (1, 2), (5, 6)])
    assert O.bfs(off, cols, 0)[0].tolist() == [0, 1, 2, -1, -1, -1, -1, -1, -1, -1]
    # complete binary tree: level = floor(log2(i + 1)) from the root
    n = 1023
    off, cols = undirected(n, [((i - 1) // 2, i) for i in range(1, n)])
    np.testing.assert_array_equal(O.bfs(off, cols, 0)[0], np.floor(np.log2(np.arange(n) + 1)).astype(int))


def test_reported_counts():
    off, cols = undirected(6, [(0, 1), (1, 2), (2, 0), (3, 4)])
    lv, _, st = O.bfs(off, cols, 0)
    assert st == {"n_reached": 3, "m_reached": 6, "num_levels": 2}


def test_check_levels_accepts_bfs_and_rejects_planted_errors():
    rng = np.random.default_rng(5)
    n = 400
    pairs = rng.integers(0, n, size=(1600, 2))
    off, cols = undirected(n, pairs)
    lv, par, _ = O.bfs(off, cols, 0, parents=True)
    assert O.check_levels(off, cols, 0, lv) == 0
    assert O.check_parents(off, cols, 0, lv, par) == 0
    reached = np.nonzero(lv > 0)[0]
    v = int(reached[len(reached) // 2])
    for delta in (-1, +1):
        bad = lv.copy()
        bad[v] += delta
        assert O.check_levels(off, cols, 0, bad) != 0
    bad = lv.copy()
    bad[v] = -1
    assert O.check_levels(off, cols, 0, bad) != 0
    # parents: wrong level, non-edge, unreached with a parent, bad root
    p2 = par.copy()
    p2[v] = v
    assert O.check_parents(off, cols, 0, lv, p2) != 0
    nonadj = [u for u in range(n) if lv[u] == lv[v] - 1 and u not in cols[off[v]:off[v + 1]]]
    if nonadj:
        p2 = par.copy()
        p2[v] = nonadj[0]
        assert O.check_parents(off, cols, 0, lv, p2) == (5, v)
    p2 = par.copy()
    p2[0] = -1
    assert O.check_parents(off, cols, 0, lv, p2) != 0


def test_scan_known_answer_and_invariants():
    # 1D Partition/misc/test_prefix.cu:91-93 scans 0..15; its printed result.
    want = [0, 0, 1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 66, 78, 91, 105]
    assert O.exclusive_scan(np.arange(16)).tolist() == want
    assert O.blelloch_scan(np.arange(16)).tolist() == want
    rng = np.random.default_rng(2)
    for n in (1, 2, 64, 1024, 4096):
        x = rng.integers(-1000, 1000, size=n).astype(np.int32)
        ex = O.exclusive_scan(x)
        assert ex[0] == 0
        np.testing.assert_array_equal(np.diff(ex.astype(np.int64)), x[:-1].astype(np.int64))
        np.testing.assert_array_equal(ex, O.blelloch_scan(x))
    # 32-bit wrap-around, as int arithmetic on the GPU
    x = np.full(4, 2**30, dtype=np.int64).astype(np.int32)
    assert O.exclusive_scan(x).tolist() == [0, 2**30, -2**31, -2**30]


def test_serialized_model():
    rng = np.random.default_rng(9)
    for trial in range(10):
        n = int(rng.integers(2, 300))
        pairs = rng.integers(0, n, size=(int(rng.integers(0, 5 * n)), 2))
        off, cols = csr_from_pairs(n, pairs)
        s = int(rng.integers(0, n))
        lv, sizes = O.serialized_bfs(off, cols, s)
        np.testing.assert_array_equal(lv, O.bfs(off, cols, s)[0])
        np.testing.assert_array_equal(sizes, np.bincount(lv[lv >= 0]))


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built (no reference sources)")


@needs_ref
@pytest.mark.parametrize("name", ["tree", "test4"])
def test_loader_matches_reference_csr(name):
    # The reference loader is 0-based and unchecked: only the files whose ids
    # fit it are compared (the 1-based files overflow it, csr.cpp:28).
    ro, ci = O.ref_load_csr(fixture_path(name))
    off, cols = O.load_edgelist(fixture_path(name))
    np.testing.assert_array_equal(ro, off)
    np.testing.assert_array_equal(ci, cols)


@needs_ref
def test_oracle_matches_reference_cpubfs():
    rng = np.random.default_rng(21)
    graphs = []
    for name in ("tree", "test4"):
        graphs.append(O.load_edgelist(fixture_path(name)))
    for name in ("test", "test2", "test3", "test_isolated_node"):
        graphs.append(O.load_edgelist(fixture_path(name), one_based=True))
    for trial in range(6):
        n = int(rng.integers(2, 3000))
        graphs.append(csr_from_pairs(n, rng.integers(0, n, size=(int(rng.integers(0, 6 * n)), 2))))
    for off, cols in graphs:
        n = off.size - 1
        for s in sorted({0, n - 1, int(rng.integers(0, n))}):
            lv_ref, _, _, it = O.ref_cpu_bfs(off, cols, s)
            lv, _, st = O.bfs(off, cols, s)
            np.testing.assert_array_equal(lv, lv_ref)
            assert st["num_levels"] == it
