"""The served workload at its real sizes: 8 graphs resident in HBM (R-MAT
scales 16/18/20/22/24 and road-like grids of 1M/4M/16M vertices, the bench's
served mix), a Poisson stream of (graph, source) queries well above the plain
loop's capacity, continuous batching on 4 lanes under a device-memory budget
small enough that admission binds (a 64-row result at scale 24 alone exceeds
it, so batches shrink and workspaces are given back).  Every served result is
compared with the same query run alone through gb_bfs: results are exact,
whatever batch, lane and neighbours a query had."""
import threading
import time

import numpy as np
import pytest

import bench
import gpu_graphs_amd as gg

pytestmark = pytest.mark.gpu

BUDGET = 4 << 30
QUERIES = 320
RATE_QPS = 160.0  # about 4x what the plain one-at-a-time loop sustains on this mix


def log(msg):
    print(msg, flush=True)


def test_served_mix_results_equal_alone_under_a_binding_budget(gpu):
    gs, dgs, cands, names = [], [], [], []
    for name, fn in bench.served_specs(gg):
        g = fn()
        gs.append(g)
        dgs.append(gg.DeviceGraph(g, gpu))
        cands.append(np.nonzero(np.diff(g.offsets))[0].astype(np.int32))
        names.append(name)
        log(f"served mix: {name} n={g.n} m={g.m}")
    rng = np.random.default_rng(bench.SEED_SERVE + 100)
    gi = rng.integers(len(gs), size=QUERIES)
    src = [int(rng.choice(cands[g])) for g in gi]
    gaps = rng.exponential(1.0 / RATE_QPS, size=QUERIES)

    srv = gg.Server(gpu, max_batch=64, lanes=4, memory_budget=BUDGET)
    gids = [srv.add_graph(d) for d in dgs]
    tickets = [None] * QUERIES
    results = [None] * QUERIES
    sizes = [0] * QUERIES
    errors = []
    peak = [0]

    def consumer():
        # Copies each result out and releases it as soon as it completes, in
        # any order (held results count against the budget).
        pending, nxt, finished = [], 0, 0
        while finished < QUERIES:
            while nxt < QUERIES and tickets[nxt] is not None:
                pending.append(nxt)
                nxt += 1
            still = []
            for i in pending:
                out = np.empty(gs[gi[i]].n, dtype=np.int32)
                rc, lat = srv.poll(tickets[i], out)
                if rc == gg.GB_EAGAIN:
                    still.append(i)
                    continue
                if rc != 0:
                    errors.append((i, rc))
                else:
                    results[i] = out
                    sizes[i] = lat["batch_size"]
                srv.release(tickets[i])
                finished += 1
            peak[0] = max(peak[0], srv.stats()["memory_used"])
            if len(still) == len(pending):
                time.sleep(0.0005)
            pending = still

    th = threading.Thread(target=consumer)
    th.start()
    t_next = time.perf_counter()
    for i in range(QUERIES):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        tickets[i] = srv.submit(gids[gi[i]], src[i])
    th.join(timeout=600)
    assert not th.is_alive(), "served queries did not complete"
    st = srv.stats()
    log(f"served mix stats {st} peak memory {peak[0]}")
    assert not errors, errors[:5]
    assert st["completed"] == QUERIES and st["failed"] == 0
    assert peak[0] <= BUDGET
    assert max(sizes) > 1  # queries were batched
    # The budget bound: the scale-24 and 16M-grid workspaces (about 2.7 and
    # 2.5 GB) cannot both be held, so lanes gave workspaces back.
    assert st["workspace_drops"] > 0

    for i in range(QUERIES):
        alone = dgs[gi[i]].bfs(src[i])[0]
        np.testing.assert_array_equal(results[i], alone,
                                      err_msg=f"query {i}: graph {names[gi[i]]} source {src[i]}")
    srv.close()
