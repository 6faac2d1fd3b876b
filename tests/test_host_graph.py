"""Host side of the engine (no GPU): the edge-list loader that replaces
csr::csr (1D Partition/src/csr.cpp:3-67), the CSR builder and the
generators.  Compared against the oracle loader, scipy, the reference's own
loader (oracle/_ref, when built) and structural invariants."""
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix

import gpu_graphs_amd as gg
from conftest import FIXTURES, GOLDEN, fixture_path
from oracle import oracle as O


def assert_csr_invariants(g, symmetric=False, no_loops=False):
    off, cols = g.offsets, g.cols
    n = g.n
    assert off[0] == 0 and off[-1] == g.m == cols.size
    assert (np.diff(off) >= 0).all()
    if cols.size:
        assert cols.min() >= 0 and cols.max() < n
    rows = np.repeat(np.arange(n), np.diff(off))
    key = rows.astype(np.int64) * n + cols
    assert (np.diff(key) > 0).all(), "rows must be sorted and duplicate-free"
    if no_loops:
        assert (rows != cols).all()
    if symmetric:
        a = csr_matrix((np.ones(cols.size), cols, off), shape=(n, n))
        assert (a != a.T).nnz == 0


@pytest.mark.parametrize("name", list(FIXTURES))
def test_loader_vs_oracle_loader(name):
    one_based, auto_n = FIXTURES[name]
    g = gg.csr(fixture_path(name), one_based=one_based, keep_order=True, auto_n=auto_n)
    if not auto_n:
        off, cols = O.load_edgelist(fixture_path(name), one_based=one_based)
        np.testing.assert_array_equal(g.row_offset, off)
        np.testing.assert_array_equal(g.col_idx, cols)
    s = gg.csr(fixture_path(name), one_based=one_based, auto_n=auto_n)
    assert_csr_invariants(s.graph)
    # same edge set as scipy's deduplicated CSR of the kept-order graph
    n = g.num_nodes
    a = csr_matrix((np.ones(g.num_edges), g.col_idx.copy(), g.row_offset.copy()), shape=(n, n))
    a.sum_duplicates()
    a.sort_indices()
    np.testing.assert_array_equal(s.row_offset, a.indptr)
    np.testing.assert_array_equal(s.col_idx, a.indices)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built")
@pytest.mark.parametrize("name", ["tree", "test4"])
def test_keep_order_matches_reference_csr(name):
    ro, ci = O.ref_load_csr(fixture_path(name))
    g = gg.csr(fixture_path(name), keep_order=True)
    np.testing.assert_array_equal(g.row_offset, ro)
    np.testing.assert_array_equal(g.col_idx, ci)


def test_tree_layout_and_header_hint():
    g = gg.csr(fixture_path("tree"))
    assert g.num_nodes == 8 and g.num_edges == 6
    assert g.row_offset.tolist() == [0, 0, 2, 4, 6, 6, 6, 6, 6]
    assert g.col_idx.tolist() == [2, 3, 4, 5, 6, 7]
    g = gg.csr(fixture_path("test4"))  # header says 7 edges, file has 5 pairs
    assert g.num_edges == 5


def test_loader_errors(tmp_path):
    def load(text, **kw):
        p = tmp_path / "g.txt"
        p.write_text(text)
        return gg.csr(str(p), **kw)

    with pytest.raises(gg.GBError) as e:
        load("4\n2\n0 1\n1 4\n")
    assert e.value.code == -5 and "outside" in str(e.value)
    with pytest.raises(gg.GBError) as e:
        load("4\n2\n0 1\n1 x\n")
    assert e.value.code == -2 and "malformed" in str(e.value)
    with pytest.raises(gg.GBError) as e:
        load("4\n2\n0 1\n1\n")
    assert e.value.code == -2 and "dangling" in str(e.value)
    with pytest.raises(gg.GBError) as e:
        load("")
    assert e.value.code == -2
    with pytest.raises(gg.GBError) as e:
        load("3\n1\n0 3\n", one_based=False)
    assert e.value.code == -5
    g = load("3\n1\n0 3\n", auto_n=True)
    assert g.num_nodes == 4
    with pytest.raises(gg.GBError):
        load("3\n1\n1 0\n", one_based=True)  # id 0 is -1 after the one-based shift
    g = load("# comment\n3\n2\n0 1\n% another\n1 2\n")
    assert g.num_edges == 2
    g = load("3\n0\n")
    assert g.num_nodes == 3 and g.num_edges == 0


def write_big_edgelist(path, n=60000, m=250000, seed=3, one_based=False, crlf=False, span_every=0,
                       comment_every=0, bad_line=None, bad_text="1 x"):
    """A multi-MiB edge list (parsed by several workers).  Returns the line
    numbers of the pairs.  span_every splits every k-th pair over two lines,
    as `ifstream >>` allows; bad_line replaces that line with bad_text."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m) + int(one_based)
    dst = rng.integers(0, n, m) + int(one_based)
    lines = [str(n), str(m)]
    for i in range(m):
        if comment_every and i % comment_every == 0:
            lines.append("# comment %d" % i)
        if span_every and i % span_every == 0:
            lines += [str(src[i]), str(dst[i])]
        else:
            lines.append("%d %d" % (src[i], dst[i]))
    if bad_line is not None:
        lines[bad_line - 1] = bad_text
    eol = "\r\n" if crlf else "\n"
    with open(path, "w", newline="") as f:
        f.write(eol.join(lines) + eol)
    assert os.path.getsize(path) > (1 << 20)


@pytest.mark.parametrize("variant", ["plain", "crlf_spanning_pairs", "one_based"])
def test_parallel_parser_matches_oracle_loader(tmp_path, monkeypatch, variant):
    p = str(tmp_path / "big.txt")
    one_based = variant == "one_based"
    write_big_edgelist(p, one_based=one_based, crlf=variant == "crlf_spanning_pairs",
                       span_every=7 if variant == "crlf_spanning_pairs" else 0)
    off, cols = O.load_edgelist(p, one_based=one_based)
    for threads in ("16", "7", "1"):
        monkeypatch.setenv("GB_THREADS", threads)
        g = gg.csr(p, one_based=one_based, keep_order=True)
        np.testing.assert_array_equal(g.row_offset, off)
        np.testing.assert_array_equal(g.col_idx, cols)


def test_parallel_parser_comments_and_auto_n(tmp_path, monkeypatch):
    p = str(tmp_path / "big.txt")
    write_big_edgelist(p, comment_every=500, span_every=11)
    got = {}
    for threads in ("16", "1"):
        monkeypatch.setenv("GB_THREADS", threads)
        g = gg.csr(p, auto_n=True)
        got[threads] = (g.num_nodes, g.row_offset.copy(), g.col_idx.copy())
    assert got["16"][0] == got["1"][0]
    np.testing.assert_array_equal(got["16"][1], got["1"][1])
    np.testing.assert_array_equal(got["16"][2], got["1"][2])
    assert_csr_invariants(gg.csr(p).graph)


@pytest.mark.parametrize("bad_text,code,needle", [
    ("1 x", -2, "malformed token on line 200001"),
    ("5 60000", -5, ":200001: edge (5, 60000) outside the 60000-vertex id range"),
    ("17", -2, "malformed token on line 200002"),  # odd token count: the pairing breaks there
])
def test_parallel_parser_errors_name_the_line(tmp_path, monkeypatch, bad_text, code, needle):
    p = str(tmp_path / "bad.txt")
    write_big_edgelist(p, bad_line=200001, bad_text=bad_text)
    if bad_text == "17":  # a lone id: the next line's pair is read across, ending on a dangling id
        with open(p) as f:
            text = f.read().splitlines()
        text[200001] = "3 y"
        with open(p, "w") as f:
            f.write("\n".join(text) + "\n")
    for threads in ("16", "1"):
        monkeypatch.setenv("GB_THREADS", threads)
        with pytest.raises(gg.GBError) as e:
            gg.csr(p)
        assert e.value.code == code and needle in str(e.value), str(e.value)


def test_parallel_parser_dangling_last_id(tmp_path, monkeypatch):
    p = str(tmp_path / "odd.txt")
    write_big_edgelist(p)
    with open(p, "a") as f:
        f.write("42\n")
    for threads in ("16", "1"):
        monkeypatch.setenv("GB_THREADS", threads)
        with pytest.raises(gg.GBError) as e:
            gg.csr(p)
        assert e.value.code == -2 and "dangling vertex id on line 250003" in str(e.value)


def test_symmetrize_and_flags(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("4\n5\n0 1\n1 2\n2 2\n0 1\n3 0\n")
    g = gg.csr(str(p), symmetrize=True)
    assert_csr_invariants(g.graph, symmetric=True)
    assert g.num_edges == 7  # {0-1, 1-2, 0-3} both ways + the 2-2 loop once
    g = gg.Graph.from_file(str(p), gg.GB_SYMMETRIZE | gg.GB_DROP_SELF_LOOPS)
    assert_csr_invariants(g, symmetric=True, no_loops=True)
    assert g.m == 6
    g = gg.csr(str(p), keep_order=True)
    assert g.col_idx.tolist() == [1, 1, 2, 2, 0]


def test_from_csr_and_binary_roundtrip(tmp_path):
    g = gg.Graph.rmat(10, seed=3)
    h = gg.Graph.from_csr(g.offsets, g.cols)
    np.testing.assert_array_equal(h.offsets, g.offsets)
    np.testing.assert_array_equal(h.cols, g.cols)
    path = str(tmp_path / "g.bin")
    g.save_binary(path)
    b = gg.Graph.load_binary(path)
    np.testing.assert_array_equal(b.offsets, g.offsets)
    np.testing.assert_array_equal(b.cols, g.cols)
    with open(path, "r+b") as f:
        f.truncate(100)
    with pytest.raises(gg.GBError):
        gg.Graph.load_binary(path)
    txt = str(tmp_path / "g.txt")
    g.write_edgelist(txt)
    t = gg.Graph.from_file(txt)
    np.testing.assert_array_equal(t.offsets, g.offsets)
    np.testing.assert_array_equal(t.cols, g.cols)


def test_rmat_properties_and_determinism(monkeypatch):
    g = gg.Graph.rmat(14, 16, seed=99)
    assert g.n == 2**14
    assert_csr_invariants(g, symmetric=True, no_loops=True)
    assert 0.5 * 2 * 16 * 2**14 < g.m <= 2 * 16 * 2**14
    deg = g.degrees()
    assert deg.max() > 20 * deg.mean()  # power-law hubs
    monkeypatch.setenv("GB_THREADS", "1")
    g1 = gg.Graph.rmat(14, 16, seed=99)
    np.testing.assert_array_equal(g1.offsets, g.offsets)
    np.testing.assert_array_equal(g1.cols, g.cols)
    g2 = gg.Graph.rmat(14, 16, seed=100)
    assert g2.m != g.m or not np.array_equal(g2.cols, g.cols)


def test_rmat_quadrant_probabilities():
    # Without the id permutation, the top bit of (u, v) follows the first
    # quadrant draw: P(u < n/2) = a + b = 0.76, P(v < n/2) = a + c = 0.76.
    g = gg.Graph.rmat(12, 16, seed=5, permute=False)
    off, cols = g.offsets, g.cols
    rows = np.repeat(np.arange(g.n), np.diff(off))
    half = g.n // 2
    both_low = np.mean((rows < half) & (cols < half))
    both_high = np.mean((rows >= half) & (cols >= half))
    # symmetrised: P(both low) ~ a = 0.57 and P(both high) ~ d = 0.05, less dedup
    assert 0.45 < both_low < 0.65 and 0.02 < both_high < 0.08


def test_er_workload_matches_golden_file():
    g = gg.Graph.erdos_renyi(10000, 200, 50000, seed=20260002)
    f = gg.Graph.from_file(os.path.join(GOLDEN, "er_10200.txt"))
    np.testing.assert_array_equal(g.offsets, f.offsets)
    np.testing.assert_array_equal(g.cols, f.cols)
    assert g.n == 10200 and g.m == 100000
    assert_csr_invariants(g, symmetric=True, no_loops=True)
    assert (g.degrees()[10000:] == 0).all()


@pytest.mark.parametrize("w,h", [(1, 1), (1, 7), (9, 1), (40, 30), (200, 300)])
def test_grid_spanning_tree_plus_extra(w, h):
    g = gg.Graph.grid(w, h, 0.05, seed=12)
    n = w * h
    assert_csr_invariants(g, symmetric=True, no_loops=True)
    grid_edges = h * (w - 1) + (h - 1) * w
    extra = int(round(0.05 * (grid_edges - (n - 1))))
    assert g.m == 2 * (n - 1 + extra)
    lv, _, st = O.bfs(g.offsets, g.cols, 0)
    assert st["n_reached"] == n  # spanning: connected
    off, cols = g.offsets, g.cols
    rows = np.repeat(np.arange(n), np.diff(off))
    d = np.abs(rows - cols)
    assert np.isin(d, [1, w]).all()  # only grid edges
    if n > 1:
        assert (lv.max() >= (w - 1) + (h - 1) - 0)  # at least the Manhattan diameter from a corner


def _mix64(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _rng_at(seed, stream, i):
    with np.errstate(over="ignore"):
        base = _mix64(np.uint64(seed) ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)))
        return _mix64(base + i.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15))


def test_grid_tree_is_the_minimum_spanning_tree_of_its_weights():
    # The tree must be the unique MST of the grid under the generator's edge
    # keys (random 32-bit weight, ties by edge id).  The keys are recomputed
    # here from the published SplitMix64 finaliser and the MST taken with
    # scipy (Kruskal/Prim), on exact integer ranks of the keys.
    from scipy.sparse.csgraph import minimum_spanning_tree
    w, h, seed = 37, 23, 77
    n = w * h
    eh = h * (w - 1)
    e = np.arange(eh + (h - 1) * w, dtype=np.int64)
    u = np.where(e < eh, (e // (w - 1)) * w + e % (w - 1), e - eh)
    v = np.where(e < eh, u + 1, u + w)
    with np.errstate(over="ignore"):
        key = (_rng_at(seed, 2, e) & np.uint64(0xFFFFFFFF00000000)) | e.astype(np.uint64)
    rank = np.empty(e.size, dtype=np.float64)
    rank[np.argsort(key)] = np.arange(1, e.size + 1)
    mst = minimum_spanning_tree(csr_matrix((rank, (u, v)), shape=(n, n))).tocoo()
    want = set(zip(np.minimum(mst.row, mst.col).tolist(), np.maximum(mst.row, mst.col).tolist()))
    g = gg.Graph.grid(w, h, 0.0, seed=seed)
    rows = np.repeat(np.arange(n), np.diff(g.offsets))
    got = {(a, b) for a, b in zip(rows.tolist(), g.cols.tolist()) if a < b}
    assert got == want


@pytest.mark.parametrize("make", [lambda: gg.Graph.rmat(12, seed=5), lambda: gg.Graph.grid(40, 30, 0.05, seed=2),
                                  lambda: gg.Graph.erdos_renyi(500, 20, 900, seed=4)])
def test_degree_buckets(make):
    # bucket 0 = degree 0, bucket b = degree in [2^(b-1), 2^b); ids ascending
    # within a bucket; the buckets partition the vertex set.
    g = make()
    off, verts = g.degree_buckets()
    deg = np.diff(g.offsets)
    assert off[0] == 0 and off[-1] == g.n and (np.diff(off) >= 0).all()
    assert np.array_equal(np.sort(verts), np.arange(g.n))
    for b in range(len(off) - 1):
        vb = verts[off[b]:off[b + 1]]
        assert (np.diff(vb) > 0).all()
        d = deg[vb]
        if b == 0:
            assert (d == 0).all()
        else:
            assert ((d >= 2 ** (b - 1)) & (d < 2 ** b)).all()
    # closed form of the bucket sizes from the degree histogram
    want = np.bincount(np.where(deg > 0, np.floor(np.log2(np.maximum(deg, 1))).astype(int) + 1, 0),
                       minlength=len(off) - 1)
    assert np.array_equal(np.diff(off), want)


def test_symmetry_detection(tmp_path):
    assert gg.Graph.rmat(10, seed=1).is_symmetric
    # directed path: not symmetric; the same edges both ways: symmetric (found
    # on sorted rows without the symmetrize flag)
    src, dst = np.arange(0, 9, dtype=np.int32), np.arange(1, 10, dtype=np.int32)
    assert not gg.Graph.from_edges(10, src, dst).is_symmetric
    both = gg.Graph.from_edges(10, np.concatenate([src, dst]), np.concatenate([dst, src]))
    assert both.is_symmetric
    # kept in file order: reported as not symmetric (rows are not checked)
    kept = gg.Graph.from_edges(10, np.concatenate([src, dst]), np.concatenate([dst, src]), flags=gg.GB_KEEP_ORDER)
    assert not kept.is_symmetric
    # one missing reverse edge breaks it
    s2 = np.concatenate([src, dst[:-1]])
    d2 = np.concatenate([dst, src[:-1]])
    assert not gg.Graph.from_edges(10, s2, d2).is_symmetric
