// Host-only tests of the query scheduler (gpu_graphs_amd/csrc/server.cpp)
// with an injected executor: batching, continuous admission, the memory
// budget, injected allocation / run failures, backpressure, cancellation and
// shutdown.  No GPU: the executor fills result rows with a pattern that
// depends on (graph, source, vertex), so misrouted rows are detected.
// Built and run by tests/test_server_host.py.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <thread>
#include <vector>

#include "server.hpp"

using namespace gg;

static int g_failures = 0;
#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failures;                                                           \
        }                                                                           \
    } while (0)

static int32_t pattern(int gid, int32_t source, int64_t v) { return int32_t((v * 7 + source * 13 + gid) % 101); }

struct FakeExec : Executor {
    std::map<int, int64_t> n;
    std::atomic<int> allocs{0};
    int fail_alloc_at = -1;
    int32_t fail_source = -1;
    int delay_ms = 0;
    std::atomic<int> running{0}, max_running{0}, runs{0};
    std::atomic<int64_t> live_alloc{0};
    std::vector<int> batch_sizes;
    std::mutex mu;

    int64_t vertices(int gid) override { return n.at(gid); }
    size_t workspace_bytes(int, bool) override { return 1000; }
    std::atomic<int> drops{0};
    void drop_workspace(int, int) noexcept override { ++drops; }
    void* alloc_result(size_t bytes) override {
        if (allocs++ == fail_alloc_at) fail(GB_ENOMEM, "injected allocation failure");
        live_alloc += int64_t(bytes);
        return std::malloc(bytes);
    }
    void free_result(void* p) noexcept override { std::free(p); }
    void run(int, int gid, const int32_t* sources, int k, int32_t* result, bool) override {
        int now = ++running;
        int prev = max_running.load();
        while (now > prev && !max_running.compare_exchange_weak(prev, now)) {
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            batch_sizes.push_back(k);
        }
        if (delay_ms) std::this_thread::sleep_for(std::chrono::milliseconds(delay_ms));
        ++runs;
        for (int i = 0; i < k; ++i) {
            if (sources[i] == fail_source) {
                --running;
                fail(GB_EDEVICE, "injected run failure");
            }
            for (int64_t v = 0; v < n.at(gid); ++v) result[int64_t(i) * n.at(gid) + v] = pattern(gid, sources[i], v);
        }
        --running;
    }
    void fetch(const int32_t* row, int32_t* host, int64_t count) override {
        std::memcpy(host, row, sizeof(int32_t) * count);
    }
    std::map<int, double> expect;  // expected batch ms per graph (set before the first submit)
    double expected_ms(int gid, int) override {
        auto it = expect.find(gid);
        return it == expect.end() ? -1.0 : it->second;
    }
};

struct Fixture {
    FakeExec* ex;
    std::unique_ptr<Server> srv;
    Fixture(ServerConfig cfg, std::map<int, int64_t> sizes, int delay = 0) {
        auto e = std::make_unique<FakeExec>();
        e->n = sizes;
        e->delay_ms = delay;
        ex = e.get();
        srv = std::make_unique<Server>(cfg, std::move(e));
        for (auto& kv : sizes) srv->add_graph(kv.first);
    }
};

static bool row_ok(FakeExec* ex, int gid, int32_t src, const std::vector<int32_t>& row) {
    for (int64_t v = 0; v < ex->n.at(gid); ++v)
        if (row[v] != pattern(gid, src, v)) return false;
    return true;
}

static void test_routing_and_batching() {
    ServerConfig cfg;
    cfg.lanes = 2;
    Fixture f(cfg, {{0, 1000}, {1, 3000}}, 2);
    std::vector<std::pair<uint64_t, std::pair<int, int32_t>>> qs;
    for (int i = 0; i < 300; ++i) {
        int gid = i % 3 == 0 ? 1 : 0;
        int32_t src = (i * 37) % 900;
        qs.push_back({f.srv->submit(gid, src), {gid, src}});
    }
    for (auto& q : qs) {
        std::vector<int32_t> row(f.ex->n.at(q.second.first));
        QueryLatency lat;
        CHECK(f.srv->wait(q.first, 10000, row.data(), &lat) == GB_OK);
        CHECK(row_ok(f.ex, q.second.first, q.second.second, row));
        CHECK(lat.batch_size >= 1 && lat.batch_size <= 64);
        CHECK(lat.total_ms >= lat.run_ms);
        f.srv->release(q.first);
    }
    ServerStats st = f.srv->stats();
    CHECK(st.completed == 300 && st.failed == 0);
    CHECK(st.mean_batch > 1.5);  // arrivals during 2 ms runs are batched
    CHECK(st.p99_ms >= st.p50_ms && st.qps > 0);
    CHECK(f.ex->max_running.load() <= 2);
}

static void test_continuous_admission() {
    ServerConfig cfg;
    cfg.lanes = 1;
    Fixture f(cfg, {{0, 500}}, 40);
    uint64_t first = f.srv->submit(0, 1);
    std::this_thread::sleep_for(std::chrono::milliseconds(10));  // first batch is running
    std::vector<uint64_t> later;
    for (int i = 0; i < 10; ++i) later.push_back(f.srv->submit(0, 10 + i));
    QueryLatency lat;
    CHECK(f.srv->wait(first, 5000, nullptr, &lat) == GB_OK);
    CHECK(lat.batch_size == 1);
    for (uint64_t t : later) {
        CHECK(f.srv->wait(t, 5000, nullptr, &lat) == GB_OK);
        CHECK(lat.batch_size == 10);  // all arrivals during the first run share the next batch
        f.srv->release(t);
    }
    f.srv->release(first);
}

static void test_plain_mode() {
    ServerConfig cfg;
    cfg.plain = true;
    cfg.lanes = 1;
    Fixture f(cfg, {{0, 200}}, 1);
    std::vector<uint64_t> ts;
    for (int i = 0; i < 20; ++i) ts.push_back(f.srv->submit(0, i));
    for (uint64_t t : ts) {
        QueryLatency lat;
        CHECK(f.srv->wait(t, 5000, nullptr, &lat) == GB_OK);
        CHECK(lat.batch_size == 1);
        f.srv->release(t);
    }
    CHECK(f.srv->stats().mean_batch == 1.0);
}

static void test_budget_limits_concurrency_and_held_results() {
    ServerConfig cfg;
    cfg.lanes = 3;
    cfg.max_batch = 4;
    const int64_t n = 1000;
    cfg.memory_budget = n * 4 + 1000 + 100;  // one result row + one lane workspace
    Fixture f(cfg, {{0, n}}, 20);
    uint64_t a = f.srv->submit(0, 1);
    CHECK(f.srv->wait(a, 5000, nullptr, nullptr) == GB_OK);
    // The only buffer is held by an unreleased result: a new query must wait.
    uint64_t b = f.srv->submit(0, 99);
    CHECK(f.srv->wait(b, 150, nullptr, nullptr) == GB_EAGAIN);
    f.srv->release(a);
    std::vector<int32_t> row(n);
    CHECK(f.srv->wait(b, 5000, row.data(), nullptr) == GB_OK);
    CHECK(row_ok(f.ex, 0, 99, row));
    f.srv->release(b);
    // Many queries: batches shrink to one row, never more than one in flight.
    std::vector<uint64_t> many;
    for (int i = 0; i < 24; ++i) many.push_back(f.srv->submit(0, i));
    for (uint64_t t : many) {
        QueryLatency lat;
        CHECK(f.srv->wait(t, 10000, nullptr, &lat) == GB_OK);
        CHECK(lat.batch_size == 1);
        f.srv->release(t);
    }
    CHECK(f.ex->max_running.load() == 1);
    CHECK(f.srv->stats().memory_used <= cfg.memory_budget);
}

static void test_batch_shrinks_to_budget() {
    // Room for 4 result rows and one workspace: 10 queued queries are served
    // in batches of at most 4 rows, whatever is held or pooled meanwhile.
    ServerConfig cfg;
    cfg.lanes = 1;
    const int64_t n = 1000;
    cfg.memory_budget = 4 * n * 4 + 1000 + 100;
    Fixture f(cfg, {{0, n}}, 30);
    uint64_t first = f.srv->submit(0, 1);
    std::this_thread::sleep_for(std::chrono::milliseconds(10));
    std::vector<uint64_t> later;
    for (int i = 0; i < 10; ++i) later.push_back(f.srv->submit(0, 20 + i));
    CHECK(f.srv->wait(first, 5000, nullptr, nullptr) == GB_OK);
    f.srv->release(first);
    int served = 0;
    for (size_t i = 0; i < later.size(); ++i) {
        std::vector<int32_t> row(n);
        QueryLatency lat;
        CHECK(f.srv->wait(later[i], 5000, row.data(), &lat) == GB_OK);
        CHECK(row_ok(f.ex, 0, int32_t(20 + i), row));
        CHECK(lat.batch_size >= 1 && lat.batch_size <= 4);
        served += 1;
        f.srv->release(later[i]);
    }
    CHECK(served == 10);
    CHECK(f.srv->stats().memory_used <= cfg.memory_budget);
    for (int k : f.ex->batch_sizes) CHECK(k <= 4);
}

static void test_query_that_never_fits() {
    ServerConfig cfg;
    cfg.lanes = 2;
    cfg.max_batch = 8;
    cfg.memory_budget = 8 * 1000 * 4 + 2000 + 64;
    Fixture f(cfg, {{0, 1000}, {1, 1000000}}, 1);
    uint64_t big = f.srv->submit(1, 5);
    uint64_t small = f.srv->submit(0, 5);
    CHECK(f.srv->wait(big, 5000, nullptr, nullptr) == GB_ENOMEM);
    std::vector<int32_t> row(1000);
    CHECK(f.srv->wait(small, 5000, row.data(), nullptr) == GB_OK);
    CHECK(row_ok(f.ex, 0, 5, row));
    f.srv->release(big);
    f.srv->release(small);
    uint64_t again = f.srv->submit(0, 6);  // still serving
    CHECK(f.srv->wait(again, 5000, nullptr, nullptr) == GB_OK);
    f.srv->release(again);
}

static void test_injected_failures() {
    ServerConfig cfg;
    cfg.lanes = 1;
    Fixture f(cfg, {{0, 300}}, 0);
    f.ex->fail_alloc_at = 0;
    uint64_t t0 = f.srv->submit(0, 1);
    CHECK(f.srv->wait(t0, 5000, nullptr, nullptr) == GB_ENOMEM);
    f.srv->release(t0);
    uint64_t t1 = f.srv->submit(0, 2);
    std::vector<int32_t> row(300);
    CHECK(f.srv->wait(t1, 5000, row.data(), nullptr) == GB_OK);
    CHECK(row_ok(f.ex, 0, 2, row));
    f.srv->release(t1);
    f.ex->fail_source = 7;
    f.ex->delay_ms = 30;
    uint64_t blocker = f.srv->submit(0, 3);
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    uint64_t bad = f.srv->submit(0, 7);
    uint64_t mate = f.srv->submit(0, 8);  // same batch as the failing source
    CHECK(f.srv->wait(blocker, 5000, nullptr, nullptr) == GB_OK);
    CHECK(f.srv->wait(bad, 5000, nullptr, nullptr) == GB_EDEVICE);
    CHECK(f.srv->wait(mate, 5000, nullptr, nullptr) == GB_EDEVICE);
    f.ex->fail_source = -1;
    f.ex->delay_ms = 0;
    uint64_t after = f.srv->submit(0, 9);
    CHECK(f.srv->wait(after, 5000, row.data(), nullptr) == GB_OK);
    CHECK(row_ok(f.ex, 0, 9, row));
    for (uint64_t t : {blocker, bad, mate, after}) f.srv->release(t);
    ServerStats st = f.srv->stats();
    CHECK(st.failed == 3 && st.completed == 3);
}

static void test_backpressure_and_cancel() {
    ServerConfig cfg;
    cfg.lanes = 1;
    cfg.queue_capacity = 3;
    Fixture f(cfg, {{0, 100}}, 50);
    uint64_t running = f.srv->submit(0, 0);
    std::this_thread::sleep_for(std::chrono::milliseconds(10));
    std::vector<uint64_t> q;
    for (int i = 0; i < 3; ++i) q.push_back(f.srv->submit(0, 1 + i));
    bool busy = false;
    try {
        f.srv->submit(0, 9);
    } catch (const Error& e) {
        busy = e.code == GB_EBUSY;
    }
    CHECK(busy);
    CHECK(f.srv->stats().rejected == 1);
    f.srv->release(q[1]);  // cancel a queued query
    CHECK(f.srv->drain(5000));
    CHECK(f.srv->wait(q[0], 100, nullptr, nullptr) == GB_OK);
    CHECK(f.srv->wait(q[2], 100, nullptr, nullptr) == GB_OK);
    bool unknown = false;
    try {
        f.srv->poll(q[1], nullptr, nullptr, nullptr);
    } catch (const Error& e) {
        unknown = e.code == GB_EINVAL;
    }
    CHECK(unknown);
    for (uint64_t t : {running, q[0], q[2]}) f.srv->release(t);
}

static void test_shutdown_with_pending() {
    ServerConfig cfg;
    cfg.lanes = 1;
    auto f = std::make_unique<Fixture>(cfg, std::map<int, int64_t>{{0, 100}}, 20);
    for (int i = 0; i < 50; ++i) f->srv->submit(0, i);
    f.reset();  // must not hang or leak
}

static void test_errors() {
    ServerConfig cfg;
    Fixture f(cfg, {{0, 10}}, 0);
    bool range = false, unknown_graph = false;
    try {
        f.srv->submit(0, 10);
    } catch (const Error& e) {
        range = e.code == GB_ERANGE;
    }
    try {
        f.srv->submit(3, 0);
    } catch (const Error& e) {
        unknown_graph = e.code == GB_EINVAL;
    }
    CHECK(range && unknown_graph);
    bool bad_cfg = false;
    try {
        ServerConfig c2;
        c2.max_batch = 65;
        Server s(c2, std::make_unique<FakeExec>());
    } catch (const Error& e) {
        bad_cfg = e.code == GB_EINVAL;
    }
    CHECK(bad_cfg);
}

static void test_idle_workspaces_make_room_for_other_graphs() {
    // Two graphs, one lane, room for one workspace and one row of either
    // graph: alternating queries switch the lane's workspace every time.
    ServerConfig cfg;
    cfg.lanes = 1;
    cfg.max_batch = 1;
    const int64_t n = 1000;
    cfg.memory_budget = n * 4 + 1000 + 100;
    Fixture f(cfg, {{0, n}, {1, n}}, 1);
    for (int i = 0; i < 6; ++i) {
        const int gid = i % 2;
        uint64_t t = f.srv->submit(gid, 5 + i);
        std::vector<int32_t> row(n);
        CHECK(f.srv->wait(t, 5000, row.data(), nullptr) == GB_OK);
        CHECK(row_ok(f.ex, gid, 5 + i, row));
        f.srv->release(t);
    }
    const ServerStats st = f.srv->stats();
    CHECK(st.completed == 6 && st.failed == 0);
    CHECK(st.workspace_drops == 5 && f.ex->drops.load() == 5);
    CHECK(st.memory_used <= cfg.memory_budget);
}

static void test_arrivals_for_a_long_graph_gather() {
    // Batches of graph 0 run long and may hold one of the two lanes: queries
    // arriving while one runs wait for the next batch, the other lane stays
    // free for short queries.
    ServerConfig cfg;
    cfg.lanes = 2;
    Fixture f(cfg, {{0, 100}}, 100);
    f.ex->expect = {{0, 100.0}};
    uint64_t first = f.srv->submit(0, 1);
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    std::vector<uint64_t> later;
    for (int i = 0; i < 10; ++i) later.push_back(f.srv->submit(0, 10 + i));
    QueryLatency lat;
    CHECK(f.srv->wait(first, 10000, nullptr, &lat) == GB_OK);
    CHECK(lat.batch_size == 1);
    for (uint64_t t : later) {
        CHECK(f.srv->wait(t, 10000, nullptr, &lat) == GB_OK);
        CHECK(lat.batch_size == 10);
        f.srv->release(t);
    }
    f.srv->release(first);
    CHECK(f.ex->max_running.load() == 1);
}

static void test_short_graph_uses_idle_lanes() {
    // Short batches (no expectation): a second query of a running graph
    // starts at once on an idle lane.
    ServerConfig cfg;
    cfg.lanes = 4;
    Fixture f(cfg, {{0, 100}}, 200);
    uint64_t a = f.srv->submit(0, 1);
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    uint64_t b = f.srv->submit(0, 2);
    QueryLatency la, lb;
    CHECK(f.srv->wait(a, 10000, nullptr, &la) == GB_OK);
    CHECK(f.srv->wait(b, 10000, nullptr, &lb) == GB_OK);
    CHECK(la.batch_size == 1 && lb.batch_size == 1);
    CHECK(lb.queue_ms < 120);  // did not wait for a's 200 ms batch
    f.srv->release(a);
    f.srv->release(b);
}

static void test_long_batches_leave_a_lane_for_short_ones() {
    // Three graphs with long batches and one short graph on three lanes: the
    // long graphs hold at most two lanes, so the short query starts at once
    // and the third long query waits for a long lane.
    ServerConfig cfg;
    cfg.lanes = 3;
    Fixture f(cfg, {{0, 100}, {1, 100}, {2, 100}, {3, 100}}, 250);
    f.ex->expect = {{0, 300.0}, {2, 300.0}, {3, 300.0}};
    uint64_t a = f.srv->submit(0, 1);
    uint64_t b = f.srv->submit(2, 1);
    uint64_t c = f.srv->submit(3, 1);
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    uint64_t s = f.srv->submit(1, 1);
    QueryLatency ls, lc;
    CHECK(f.srv->wait(s, 10000, nullptr, &ls) == GB_OK);
    CHECK(ls.queue_ms < 150);
    CHECK(f.srv->wait(c, 10000, nullptr, &lc) == GB_OK);
    CHECK(lc.queue_ms >= 150);
    for (uint64_t t : {a, b}) CHECK(f.srv->wait(t, 10000, nullptr, nullptr) == GB_OK);
    for (uint64_t t : {a, b, c, s}) f.srv->release(t);
    CHECK(f.ex->max_running.load() <= 3);
}

int main() {
    struct {
        const char* name;
        void (*fn)();
    } tests[] = {{"routing_and_batching", test_routing_and_batching},
                 {"continuous_admission", test_continuous_admission},
                 {"plain_mode", test_plain_mode},
                 {"budget", test_budget_limits_concurrency_and_held_results},
                 {"batch_shrinks_to_budget", test_batch_shrinks_to_budget},
                 {"never_fits", test_query_that_never_fits},
                 {"injected_failures", test_injected_failures},
                 {"backpressure_and_cancel", test_backpressure_and_cancel},
                 {"shutdown_with_pending", test_shutdown_with_pending},
                 {"errors", test_errors},
                 {"long_graph_arrivals_gather", test_arrivals_for_a_long_graph_gather},
                 {"short_graph_uses_idle_lanes", test_short_graph_uses_idle_lanes},
                 {"long_batches_leave_a_lane", test_long_batches_leave_a_lane_for_short_ones},
                 {"idle_workspaces_make_room", test_idle_workspaces_make_room_for_other_graphs}};
    for (auto& t : tests) {
        const int before = g_failures;
        t.fn();
        std::printf("%-28s %s\n", t.name, g_failures == before ? "ok" : "FAILED");
        std::fflush(stdout);
    }
    return g_failures ? 1 : 0;
}
