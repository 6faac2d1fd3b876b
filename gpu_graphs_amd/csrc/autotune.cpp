// Autotuner for the kernel shapes and thresholds of gb_bfs and gb_bfs_batch.
//
// No counterpart in the reference, whose kernel shape is a compile-time
// constant (BLOCK_SIZE 1024, 1D Partition/include/gpuBFS_kernels.cuh) and
// which has no thresholds (top-down only).  Coordinate descent on one graph
// handle: starting from its current configuration, each parameter in turn is
// set to each of its candidate values with the others fixed, the device time
// per call is measured (HIP events on the runner's stream; the best of `reps`
// passes after one warm-up pass, so that graph capture and the relearning of
// the launch window do not count), and the fastest value is kept.  The
// winning configuration stays on the handle.  Results never depend on the
// configuration: every candidate is a complete, exact search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <limits>
#include <vector>

#include "device_graph.hpp"
#include "handles.hpp"

namespace gg {

namespace {

// A kernel shape (workgroup size, edges per thread) as one value.
constexpr int64_t shape(int block, int k) { return int64_t(block) * 1000 + k; }

double best_of(const std::function<double()>& pass, int reps) {
    pass();  // warm-up: capture, workspace, launch-window learning
    double best = std::numeric_limits<double>::max();
    for (int r = 0; r < reps; ++r) best = std::min(best, pass());
    return best;
}

struct Knob {
    std::vector<int64_t> values;
    std::function<void(int64_t)> set;
    std::function<int64_t()> get;
};

// Tries every value of every knob in turn, keeping the fastest.
void descend(std::vector<Knob>& knobs, const std::function<double()>& measure, double& best, int& tried) {
    for (Knob& kb : knobs) {
        const int64_t keep = kb.get();
        int64_t best_v = keep;
        for (int64_t v : kb.values) {
            if (v == keep) continue;
            kb.set(v);
            const double t = measure();
            ++tried;
            if (t < best) {
                best = t;
                best_v = v;
            }
        }
        kb.set(best_v);
    }
}

void check_sources(const DeviceGraph& g, const int32_t* sources, int count) {
    if (!sources || count < 1 || count > 64) fail(GB_EINVAL, "autotune takes 1 to 64 sources");
    for (int i = 0; i < count; ++i)
        if (sources[i] < 0 || sources[i] >= g.n) fail(GB_ERANGE, "autotune source outside the graph");
}

struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) { GG_HIP(hipMalloc(&p, bytes)); }
    ~DeviceBuffer() { (void)hipFree(p); }
};

}  // namespace

void autotune_bfs(DeviceGraph& g, const int32_t* sources, int count, int reps, gb_tune_result* out) {
    check_sources(g, sources, count);
    reps = std::max(reps, 1);
    Runner& r = *g.runner;
    DeviceGuard guard(g.device);
    DeviceBuffer levels(sizeof(int32_t) * g.n);
    auto pass = [&]() {
        double t = 0;
        for (int i = 0; i < count; ++i) {
            r.bfs(sources[i], static_cast<int32_t*>(levels.p), nullptr, nullptr);
            t += r.last_ms;
        }
        return t / count;
    };
    auto measure = [&]() { return best_of(pass, reps); };
    BfsConfig& c = g.cfg;
    std::vector<Knob> knobs;
    knobs.push_back({{shape(256, 4), shape(256, 2), shape(256, 8), shape(512, 2), shape(1024, 1), shape(128, 8)},
                     [&](int64_t v) {
                         c.block = int(v / 1000);
                         c.edges_per_thread = int(v % 1000);
                     },
                     [&]() { return shape(c.block, c.edges_per_thread); }});
    if (g.symmetric && c.direction_opt) {
        auto window = [&]() { r.reset_bottom_up_window(); };
        knobs.push_back({{8, 14, 32, 64},
                         [&, window](int64_t v) {
                             c.alpha = int(v);
                             window();
                         },
                         [&]() { return int64_t(c.alpha); }});
        knobs.push_back({{8, 24, 64, 256},
                         [&, window](int64_t v) {
                             c.beta = int(v);
                             window();
                         },
                         [&]() { return int64_t(c.beta); }});
    }
    const int64_t dense0 = g.default_dense_edges();
    knobs.push_back({{dense0 / 4, dense0, dense0 * 4}, [&](int64_t v) { c.dense_edges = v; },
                     [&]() { return c.dense_edges; }});
    gb_tune_result res{};
    res.default_ms = measure();
    res.best_ms = res.default_ms;
    res.tried = 1;
    descend(knobs, measure, res.best_ms, res.tried);
    res.block = c.block;
    res.edges_per_thread = c.edges_per_thread;
    res.alpha = c.alpha;
    res.beta = c.beta;
    res.low_degree = 0;
    res.frontier_edges = c.dense_edges;
    if (out) *out = res;
}

void autotune_batch(DeviceGraph& g, const int32_t* sources, int k, int reps, gb_tune_result* out) {
    check_sources(g, sources, k);
    reps = std::max(reps, 1);
    Runner& r = *g.runner;
    DeviceGuard guard(g.device);
    DeviceBuffer levels(sizeof(int32_t) * g.n * k);
    auto pass = [&]() {
        r.bfs_batch(sources, k, static_cast<int32_t*>(levels.p), nullptr);
        return r.last_ms;
    };
    auto measure = [&]() { return best_of(pass, reps); };
    MsConfig& c = g.ms_cfg;
    if (c.sparse_edges <= 0) c.sparse_edges = g.default_ms_sparse_edges();
    std::vector<Knob> knobs;
    knobs.push_back({{shape(256, 2), shape(256, 1), shape(256, 4), shape(512, 2), shape(1024, 1)},
                     [&](int64_t v) {
                         c.block = int(v / 1000);
                         c.edges_per_thread = int(v % 1000);
                     },
                     [&]() { return shape(c.block, c.edges_per_thread); }});
    if (g.symmetric && c.direction_opt) {
        knobs.push_back({{1, 2, 4, 8}, [&](int64_t v) { c.alpha = int(v); }, [&]() { return int64_t(c.alpha); }});
        knobs.push_back({{0, 8, 32, 128}, [&](int64_t v) { c.beta = int(v); }, [&]() { return int64_t(c.beta); }});
        knobs.push_back({{16, 32, 64}, [&](int64_t v) { c.low_degree = int(v); },
                         [&]() { return int64_t(c.low_degree); }});
    }
    const int64_t sparse0 = g.default_ms_sparse_edges();
    knobs.push_back({{std::max<int64_t>(1, sparse0 / 8), sparse0, sparse0 * 8},
                     [&](int64_t v) { c.sparse_edges = v; }, [&]() { return c.sparse_edges; }});
    gb_tune_result res{};
    res.default_ms = measure();
    res.best_ms = res.default_ms;
    res.tried = 1;
    descend(knobs, measure, res.best_ms, res.tried);
    res.block = c.block;
    res.edges_per_thread = c.edges_per_thread;
    res.alpha = c.alpha;
    res.beta = c.beta;
    res.low_degree = c.low_degree;
    res.frontier_edges = c.sparse_edges;
    if (out) *out = res;
}

}  // namespace gg
