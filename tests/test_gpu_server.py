"""GPU tests of the continuous-batching query server: every served result is
bit-exact with the same query run alone (single-source gb_bfs), whether it
ran in a batch, concurrently with batches on other graphs and lanes, under a
tight memory budget, or in plain one-at-a-time mode."""
import threading
import time

import numpy as np
import pytest

import gpu_graphs_amd as gg

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def graphs(gpu):
    gs = [gg.Graph.rmat(14, seed=5), gg.Graph.grid(150, 120, 0.05, seed=6), gg.Graph.rmat(16, seed=7),
          gg.Graph.erdos_renyi(3000, 50, 9000, seed=8)]
    dgs = [gg.DeviceGraph(g, gpu) for g in gs]
    return gs, dgs


def random_queries(gs, count, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(count):
        gi = int(rng.integers(len(gs)))
        deg = np.diff(gs[gi].offsets)
        qs.append((gi, int(rng.choice(np.nonzero(deg)[0]))))
    return qs


def serve_and_check(srv, gids, gs, dgs, queries, spread_s=0.0):
    tickets = []
    t0 = time.perf_counter()
    for i, (gi, s) in enumerate(queries):
        if spread_s:
            target = t0 + spread_s * i / len(queries)
            while time.perf_counter() < target:
                pass
        tickets.append(srv.submit(gids[gi], s))
    sizes = []
    for (gi, s), t in zip(queries, tickets):
        out = np.empty(gs[gi].n, dtype=np.int32)
        rc, lat = srv.wait(t, out, timeout_ms=60000)
        assert rc == 0, rc
        alone = dgs[gi].bfs(s)[0]
        np.testing.assert_array_equal(out, alone, err_msg=f"graph {gi} source {s}")
        sizes.append(lat["batch_size"])
        srv.release(t)
    return sizes


def test_served_results_equal_alone(graphs):
    gs, dgs = graphs
    srv = gg.Server(lanes=4, max_batch=64)
    gids = [srv.add_graph(d) for d in dgs]
    qs = random_queries(gs, 400, 1)
    sizes = serve_and_check(srv, gids, gs, dgs, qs)
    st = srv.stats()
    log(f"served {st}")
    assert st["completed"] == 400 and st["failed"] == 0
    assert max(sizes) > 1  # queries were batched
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_arrivals_during_execution_join_later_batches(graphs):
    gs, dgs = graphs
    srv = gg.Server(lanes=2, max_batch=32)
    gids = [srv.add_graph(d) for d in dgs]
    qs = random_queries(gs, 300, 2)
    sizes = serve_and_check(srv, gids, gs, dgs, qs, spread_s=0.3)
    st = srv.stats()
    log(f"spread arrivals: batches {st['batches']} mean {st['mean_batch']:.2f}")
    assert st["batches"] > 2 and st["completed"] == 300
    assert max(sizes) <= 32


def test_tight_budget_still_serves_everything(graphs):
    gs, dgs = graphs
    # Room for about two result buffers of the largest graph plus workspaces.
    n_max = max(g.n for g in gs)
    budget = 2 * 64 * n_max * 4 + 40 * n_max * 8
    srv = gg.Server(lanes=4, max_batch=64, memory_budget=budget)
    gids = [srv.add_graph(d) for d in dgs]
    qs = random_queries(gs, 200, 3)
    serve_and_check(srv, gids, gs, dgs, qs)
    st = srv.stats()
    assert st["completed"] == 200 and st["memory_used"] <= budget
    srv.close()


def test_budget_too_small_fails_cleanly_and_keeps_serving(graphs):
    gs, dgs = graphs
    small = min(range(len(gs)), key=lambda i: gs[i].n)
    big = max(range(len(gs)), key=lambda i: gs[i].n)
    budget = 64 * gs[small].n * 4 + 60 * gs[small].n * 8 + gs[small].m * 2
    srv = gg.Server(lanes=1, max_batch=64, memory_budget=budget)
    gids = [srv.add_graph(d) for d in dgs]
    t_big = srv.submit(gids[big], 0)
    rc, _ = srv.wait(t_big, timeout_ms=30000)
    assert rc == -3  # GB_ENOMEM: its batch can never fit
    srv.release(t_big)
    deg = np.diff(gs[small].offsets)
    s = int(np.nonzero(deg)[0][0])
    t = srv.submit(gids[small], s)
    out = np.empty(gs[small].n, dtype=np.int32)
    assert srv.wait(t, out, timeout_ms=30000)[0] == 0
    np.testing.assert_array_equal(out, dgs[small].bfs(s)[0])
    srv.release(t)
    srv.close()


def test_plain_mode_one_at_a_time(graphs):
    gs, dgs = graphs
    srv = gg.Server(lanes=1, plain=True)
    gids = [srv.add_graph(d) for d in dgs]
    sizes = serve_and_check(srv, gids, gs, dgs, random_queries(gs, 60, 4))
    assert set(sizes) == {1}
    srv.close()


def test_concurrent_clients(graphs):
    # Four client threads submit and wait concurrently (gb_server_* calls are
    # thread-safe).  The alone-references are computed afterwards on one
    # thread: a DeviceGraph handle itself is not thread-safe.
    gs, dgs = graphs
    srv = gg.Server(lanes=4, max_batch=64)
    gids = [srv.add_graph(d) for d in dgs]
    results, errors = [], []
    lock = threading.Lock()

    def client(seed):
        try:
            qs = random_queries(gs, 50, 100 + seed)
            tickets = [srv.submit(gids[gi], s) for gi, s in qs]
            for (gi, s), t in zip(qs, tickets):
                out = np.empty(gs[gi].n, dtype=np.int32)
                assert srv.wait(t, out, timeout_ms=60000)[0] == 0
                srv.release(t)
                with lock:
                    results.append((gi, s, out))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    threads = [threading.Thread(target=client, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert len(results) == 200 and srv.stats()["completed"] == 200
    for gi, s, out in results:
        np.testing.assert_array_equal(out, dgs[gi].bfs(s)[0])
    srv.close()


def test_served_large_graphs_on_concurrent_lanes(gpu):
    # Graphs with more than 1024 compaction chunks (8.4M and 4.8M vertices;
    # the grid is ~5500 levels deep) served by 4 lanes in small batches, so
    # several batches' level kernels share the device at once.
    gs = [gg.Graph.rmat(23, edge_factor=2, seed=31), gg.Graph.grid(2200, 2200, 0.05, seed=32)]
    dgs = [gg.DeviceGraph(g, gpu) for g in gs]
    log(f"large mix: n={[g.n for g in gs]} m={[g.m for g in gs]}")
    srv = gg.Server(lanes=4, max_batch=8)
    gids = [srv.add_graph(d) for d in dgs]
    qs = random_queries(gs, 64, 5)
    tickets = [srv.submit(gids[gi], s) for gi, s in qs]
    outs = []
    for (gi, s), t in zip(qs, tickets):
        out = np.empty(gs[gi].n, dtype=np.int32)
        rc, lat = srv.wait(t, out, timeout_ms=120000)
        assert rc == 0, (rc, gi, s)
        outs.append(out)
        srv.release(t)
    st = srv.stats()
    log(f"large mix served: {st}")
    assert st["completed"] == 64 and st["failed"] == 0 and st["batches"] >= 8
    srv.close()
    for (gi, s), out in zip(qs, outs):
        np.testing.assert_array_equal(out, dgs[gi].bfs(s)[0], err_msg=f"graph {gi} source {s}")


def test_server_errors(graphs):
    gs, dgs = graphs
    srv = gg.Server(lanes=1)
    gid = srv.add_graph(dgs[0])
    with pytest.raises(gg.GBError) as e:
        srv.submit(gid, gs[0].n)
    assert e.value.code == -5
    with pytest.raises(gg.GBError):
        srv.submit(gid + 7, 0)
    with pytest.raises(gg.GBError):
        gg.Server(max_batch=65)
    srv.close()
