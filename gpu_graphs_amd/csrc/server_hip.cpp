// HIP executor of the query server and its C ABI (gb_server_*).
//
// Lane i owns one stream, created with the others when the server starts,
// and one Runner (workspaces) per graph it has served, launching on the
// lane's stream; so batches on different lanes -- on the same graph or on
// different graphs -- run concurrently on the device.  (A stream per runner
// spread the lanes over the device's few hardware queues at random: two
// lanes that shared a queue ran one after the other, and a 0.2 ms R-MAT
// search waited behind a 400 ms grid batch.)  Result buffers are plain
// device allocations; the server pools and budgets them.
//
// A batch of k queries runs either as one multi-source BFS or as k
// single-source BFS, whichever the measured per-graph costs say is cheaper:
// a multi-source level scans per-vertex masks, so on deep, thin graphs a
// small batch can cost more than its queries one by one.  The costs are
// running averages of the device time of past runs on that graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "device_graph.hpp"
#include "handles.hpp"
#include "server.hpp"

namespace gg {

class HipExecutor : public Executor {
   public:
    HipExecutor(int device, int lanes) : device_(device), runners_(lanes), lane_streams_(lanes, nullptr) {
        DeviceGuard guard(device_);
        try {
            GG_HIP(hipStreamCreateWithFlags(&fetch_stream_, hipStreamNonBlocking));
            for (hipStream_t& s : lane_streams_) GG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        } catch (...) {
            release_streams();
            throw;
        }
    }
    ~HipExecutor() override {
        (void)hipSetDevice(device_);
        runners_.clear();  // each runner synchronises its lane's stream first
        release_streams();
    }

    int add(DeviceGraph* g) {
        if (g->device != device_) fail(GB_EINVAL, "graph lives on another device than the server");
        std::lock_guard<std::mutex> lk(mu_);
        graphs_.push_back(g);
        return static_cast<int>(graphs_.size()) - 1;
    }

    int64_t vertices(int gid) override { return graph(gid)->n; }

    size_t workspace_bytes(int gid, bool plain) override {
        DeviceGraph* g = graph(gid);
        // single-source workspace: two queues, bitmaps, tile index, staging
        const size_t single = size_t(g->n) * (2 * (8 + 4 + 4) + 8) + size_t(g->m / 64 + 64) * 8 + (size_t(g->n) / 4) + size_t(g->n) / 8;
        return plain ? single : single + batch_bytes_estimate(g->n, g->m, 64, false);
    }

    void drop_workspace(int lane, int gid) noexcept override {
        int prev = -1;
        const bool restore = hipGetDevice(&prev) == hipSuccess && prev != device_;
        runners_[lane].erase(gid);  // ~Runner sets the device and frees its stream and buffers
        if (restore) (void)hipSetDevice(prev);
    }

    void* alloc_result(size_t bytes) override {
        DeviceGuard guard(device_);
        void* p = nullptr;
        GG_HIP(hipMalloc(&p, bytes ? bytes : 1));
        return p;
    }

    // Called from the scheduler thread and destructors: never throws, also
    // after a device error has made every HIP call fail.
    void free_result(void* p) noexcept override {
        if (!p) return;
        int prev = -1;
        const bool restore = hipGetDevice(&prev) == hipSuccess && prev != device_;
        if (restore) (void)hipSetDevice(device_);
        (void)hipFree(p);
        if (restore) (void)hipSetDevice(prev);
    }

    void run(int lane, int gid, const int32_t* sources, int k, int32_t* result, bool plain) override {
        DeviceGraph* g = graph(gid);
        std::unique_ptr<Runner>& r = runners_[lane][gid];  // only this lane's thread touches runners_[lane]
        if (!r) r = std::make_unique<Runner>(*g, lane_streams_[lane]);
        if (plain || prefer_single(gid, k)) {
            double total = 0;
            for (int i = 0; i < k; ++i) {
                r->bfs(sources[i], result + int64_t(i) * g->n, nullptr, nullptr);
                total += r->last_ms;
            }
            record(gid, true, 1, total / k);
        } else {
            r->bfs_batch(sources, k, result, nullptr);
            record(gid, false, k, r->last_ms);
        }
    }

    // The cheaper of the two ways run() can take, from the running averages.
    double expected_ms(int gid, int k) override {
        std::lock_guard<std::mutex> lk(cost_mu_);
        const Cost& c = cost_[gid];
        const double single = c.single_ms >= 0 ? k * c.single_ms : -1.0;
        const double batch = k == 1 ? -1.0 : batch_estimate(c, k);
        if (batch < 0) return single;
        return single < 0 ? batch : std::min(single, batch);
    }

    void fetch(const int32_t* row, int32_t* host, int64_t n) override {
        DeviceGuard guard(device_);
        std::lock_guard<std::mutex> lk(fetch_mu_);
        GG_HIP(hipMemcpyAsync(host, row, sizeof(int32_t) * n, hipMemcpyDeviceToHost, fetch_stream_));
        GG_HIP(hipStreamSynchronize(fetch_stream_));
    }

    int device() const { return device_; }

   private:
    // Running averages of device time per graph: per query run alone, and per
    // multi-source batch by size class (k in (2^(c-1), 2^c], c = 0..6): on
    // deep graphs a batch of 64 costs several times a batch of 4.
    static constexpr int kClasses = 7;
    struct Cost {
        double single_ms = -1;  // per query
        double batch_ms[kClasses] = {-1, -1, -1, -1, -1, -1, -1};
    };
    static int size_class(int k) {
        int c = 0;
        while ((1 << c) < k && c < kClasses - 1) ++c;
        return c;
    }
    // Batch cost of size k: its class's average, else the nearest class
    // measured (classes below scaled up by the size ratio, an upper bound:
    // a batch never costs more than its queries one at a time).
    static double batch_estimate(const Cost& c, int k) {
        const int kc = size_class(k);
        if (c.batch_ms[kc] >= 0) return c.batch_ms[kc];
        for (int d = 1; d < kClasses; ++d) {
            if (kc + d < kClasses && c.batch_ms[kc + d] >= 0) return c.batch_ms[kc + d];
            if (kc - d >= 0 && c.batch_ms[kc - d] >= 0) return c.batch_ms[kc - d] * double(1 << d);
        }
        return -1;
    }

    // Unknown costs are learned first: a lone query runs single-source, a
    // batch of two or more runs multi-source.
    bool prefer_single(int gid, int k) {
        std::lock_guard<std::mutex> lk(cost_mu_);
        const Cost& c = cost_[gid];
        if (k == 1 && c.single_ms < 0) return true;
        const double b = batch_estimate(c, k);
        if (b < 0) return k == 1;
        if (c.single_ms < 0) return false;
        return k * c.single_ms <= b;
    }

    void record(int gid, bool single, int k, double ms) {
        std::lock_guard<std::mutex> lk(cost_mu_);
        double& avg = single ? cost_[gid].single_ms : cost_[gid].batch_ms[size_class(k)];
        avg = avg < 0 ? ms : 0.7 * avg + 0.3 * ms;
    }

    DeviceGraph* graph(int gid) {
        std::lock_guard<std::mutex> lk(mu_);
        if (gid < 0 || gid >= static_cast<int>(graphs_.size())) fail(GB_EINVAL, "unknown graph id");
        return graphs_[gid];
    }

    void release_streams() noexcept {
        for (hipStream_t& s : lane_streams_) {
            if (s) (void)hipStreamDestroy(s);
            s = nullptr;
        }
        if (fetch_stream_) (void)hipStreamDestroy(fetch_stream_);
        fetch_stream_ = nullptr;
    }

    int device_;
    std::mutex mu_, fetch_mu_, cost_mu_;
    std::map<int, Cost> cost_;
    std::vector<DeviceGraph*> graphs_;
    std::vector<std::map<int, std::unique_ptr<Runner>>> runners_;
    std::vector<hipStream_t> lane_streams_;  // one per lane, shared by its runners
    hipStream_t fetch_stream_ = nullptr;
};

}  // namespace gg

struct gb_server {
    gg::HipExecutor* exec = nullptr;  // owned by srv
    std::unique_ptr<gg::Server> srv;
};

namespace {

template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return GB_OK;
    } catch (const gg::Error& e) {
        gg::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        gg::set_last_error("host allocation failed");
        return GB_ENOMEM;
    } catch (const std::exception& e) {
        gg::set_last_error(e.what());
        return GB_EINVAL;
    }
}

}  // namespace

extern "C" {

int gb_server_create(const gb_server_cfg* cfg, gb_server** out) {
    return guarded([&] {
        if (!cfg || !out) gg::fail(GB_EINVAL, "null argument");
        *out = nullptr;
        gg::ServerConfig sc;
        sc.max_batch = cfg->max_batch > 0 ? cfg->max_batch : 64;
        sc.lanes = cfg->lanes > 0 ? cfg->lanes : 4;
        sc.plain = cfg->plain != 0;
        sc.queue_capacity = cfg->queue_capacity > 0 ? cfg->queue_capacity : (int64_t(1) << 20);
        int64_t budget = cfg->memory_budget;
        if (budget <= 0) {  // default: half of the device memory free now
            gg::DeviceGuard guard(cfg->device);
            size_t free_b = 0, total_b = 0;
            GG_HIP(hipMemGetInfo(&free_b, &total_b));
            budget = static_cast<int64_t>(free_b / 2);
        }
        sc.memory_budget = budget;
        auto h = std::make_unique<gb_server>();
        auto exec = std::make_unique<gg::HipExecutor>(cfg->device, sc.lanes);
        h->exec = exec.get();
        h->srv = std::make_unique<gg::Server>(sc, std::move(exec));
        *out = h.release();
    });
}

int gb_server_add_graph(gb_server* s, gb_dgraph* dg, int* gid) {
    return guarded([&] {
        if (!s || !dg || !gid) gg::fail(GB_EINVAL, "null argument");
        const int id = s->exec->add(dg->d.get());
        *gid = s->srv->add_graph(id);
    });
}

int gb_server_submit(gb_server* s, int gid, int32_t source, uint64_t* ticket) {
    return guarded([&] {
        if (!s || !ticket) gg::fail(GB_EINVAL, "null argument");
        *ticket = s->srv->submit(gid, source);
    });
}

static void to_latency(const gg::QueryLatency& q, gb_latency* lat) {
    if (!lat) return;
    lat->queue_ms = q.queue_ms;
    lat->run_ms = q.run_ms;
    lat->total_ms = q.total_ms;
    lat->batch_size = q.batch_size;
}

int gb_server_poll(gb_server* s, uint64_t ticket, int32_t* levels, gb_latency* lat) {
    int st = GB_OK;
    int rc = guarded([&] {
        if (!s) gg::fail(GB_EINVAL, "null server");
        gg::QueryLatency q;
        st = s->srv->poll(ticket, levels, &q, nullptr);
        to_latency(q, lat);
    });
    return rc != GB_OK ? rc : st;
}

int gb_server_wait(gb_server* s, uint64_t ticket, int timeout_ms, int32_t* levels, gb_latency* lat) {
    int st = GB_OK;
    int rc = guarded([&] {
        if (!s) gg::fail(GB_EINVAL, "null server");
        gg::QueryLatency q;
        st = s->srv->wait(ticket, timeout_ms, levels, &q);
        to_latency(q, lat);
    });
    return rc != GB_OK ? rc : st;
}

int gb_server_result_device(gb_server* s, uint64_t ticket, const int32_t** d_levels) {
    int st = GB_OK;
    int rc = guarded([&] {
        if (!s || !d_levels) gg::fail(GB_EINVAL, "null argument");
        st = s->srv->poll(ticket, nullptr, nullptr, d_levels);
    });
    return rc != GB_OK ? rc : st;
}

int gb_server_release(gb_server* s, uint64_t ticket) {
    return guarded([&] {
        if (!s) gg::fail(GB_EINVAL, "null server");
        s->srv->release(ticket);
    });
}

int gb_server_drain(gb_server* s, int timeout_ms) {
    int st = GB_OK;
    int rc = guarded([&] {
        if (!s) gg::fail(GB_EINVAL, "null server");
        st = s->srv->drain(timeout_ms) ? GB_OK : GB_EAGAIN;
    });
    return rc != GB_OK ? rc : st;
}

int gb_server_get_stats(gb_server* s, gb_server_stats* out) {
    return guarded([&] {
        if (!s || !out) gg::fail(GB_EINVAL, "null argument");
        const gg::ServerStats st = s->srv->stats();
        out->submitted = st.submitted;
        out->completed = st.completed;
        out->failed = st.failed;
        out->rejected = st.rejected;
        out->batches = st.batches;
        out->mean_batch = st.mean_batch;
        out->p50_ms = st.p50_ms;
        out->p99_ms = st.p99_ms;
        out->max_ms = st.max_ms;
        out->qps = st.qps;
        out->memory_used = st.memory_used;
        out->memory_budget = st.memory_budget;
        out->workspace_drops = st.workspace_drops;
    });
}

int gb_server_reset_stats(gb_server* s) {
    return guarded([&] {
        if (!s) gg::fail(GB_EINVAL, "null server");
        s->srv->reset_stats();
    });
}

void gb_server_free(gb_server* s) { delete s; }

}  // extern "C"
