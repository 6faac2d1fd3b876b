// Continuous-batching BFS query server (host side, no HIP dependency).
//
// No counterpart in the reference, which answers exactly one (graph, source)
// per process run (1D Partition/main.cu:10-82).  Queries (graph id, source)
// arrive at any time; a scheduler thread forms batches of up to max_batch
// queries on one graph and hands them to idle lanes.  Each lane executes one
// batch at a time on its own GPU stream, so batches on different lanes run
// concurrently on the device, and queries that arrive while lanes are busy
// join the next batch as soon as a lane and memory free up.  Graphs whose
// batches are expected to run long (deep graphs) hold at most lanes - 1 lanes
// together, so short queries always find a lane; once they hold that many,
// their arrivals gather for their next batch.  Among the graphs that may
// start a batch, the one whose oldest query has waited longest goes first.
// Admission is by
// a device-memory budget: a batch of k queries needs a result buffer of k
// rows of n int32 (rounded up to a power of two, pooled per graph and size)
// and, the first time a lane serves a graph, that lane's workspace.  When
// the budget cannot hold the whole batch it shrinks to what fits.  Results
// stay in device memory until the client releases them.
//
// The device work is behind the Executor interface; the HIP executor runs the
// multi-source BFS (or, in plain mode, one single-source BFS per query), and
// tests inject a host executor to exercise the scheduler without a GPU.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace gg {

struct ServerConfig {
    int max_batch = 64;            // queries per batch (1..64)
    int lanes = 4;                 // concurrent batches (streams)
    bool plain = false;            // one single-source BFS per query, no batching
    int64_t memory_budget = 0;     // bytes of device memory for workspaces + results
    int64_t queue_capacity = 1 << 20;
    double long_batch_ms = 10.0;   // batches expected to run longer hold at most lanes - 1 lanes
};

class Executor {
   public:
    virtual ~Executor() = default;
    virtual int64_t vertices(int gid) = 0;
    // Workspace a lane needs to serve graph gid (charged once per lane/graph).
    virtual size_t workspace_bytes(int gid, bool plain) = 0;
    // Frees the workspace `lane` holds for graph gid.  Called only while the
    // lane is idle; must not throw.
    virtual void drop_workspace(int lane, int gid) noexcept = 0;
    // Result buffers of `bytes`; alloc throws gg::Error (GB_ENOMEM) on failure.
    virtual void* alloc_result(size_t bytes) = 0;
    virtual void free_result(void* p) noexcept = 0;  // must not throw (destructors, scheduler)
    // Runs one batch on `lane`: row i of `result` (n int32) = levels from
    // sources[i].  Blocking; throws gg::Error on failure.
    virtual void run(int lane, int gid, const int32_t* sources, int k, int32_t* result, bool plain) = 0;
    // Copies n int32 of a result row to the host.
    virtual void fetch(const int32_t* row, int32_t* host, int64_t n) = 0;
    // Expected run time of a batch of k queries on graph gid in ms (learned
    // from earlier runs), or a negative value when not known yet.
    virtual double expected_ms(int gid, int k) {
        (void)gid;
        (void)k;
        return -1.0;
    }
};

struct QueryLatency {
    double queue_ms = 0;  // arrival -> batch start
    double run_ms = 0;    // batch start -> batch done
    double total_ms = 0;  // arrival -> done
    int batch_size = 0;
};

struct ServerStats {
    int64_t submitted = 0, completed = 0, failed = 0, rejected = 0, batches = 0;
    double mean_batch = 0, p50_ms = 0, p99_ms = 0, max_ms = 0, qps = 0;
    int64_t memory_used = 0, memory_budget = 0;
    int64_t workspace_drops = 0;
};

class Server {
   public:
    using Clock = std::chrono::steady_clock;

    Server(const ServerConfig& cfg, std::unique_ptr<Executor> exec);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int add_graph(int gid_hint = -1);  // registers executor graph `gid_hint` (or the next id)
    uint64_t submit(int gid, int32_t source);
    // GB_OK when done (copies the row if host != nullptr), GB_EAGAIN while
    // pending, or the query's failure status.
    int poll(uint64_t ticket, int32_t* host, QueryLatency* lat, const int32_t** device_row);
    int wait(uint64_t ticket, int timeout_ms, int32_t* host, QueryLatency* lat);
    void release(uint64_t ticket);
    bool drain(int timeout_ms);
    ServerStats stats();
    void reset_stats();
    const char* query_error(uint64_t ticket);

   private:
    struct Batch;
    struct Query {
        uint64_t ticket;
        int gid;
        int32_t source;
        Clock::time_point arrival, start, done;
        int status = GB_EAGAIN;
        std::string error;
        std::shared_ptr<Batch> batch;
        int row = 0;
        bool released = false;
    };
    struct Buffer {
        void* ptr = nullptr;
        size_t bytes = 0;
        int gid = -1;
        int rows = 0;
    };
    struct Batch {
        int gid = -1;
        int lane = -1;
        std::vector<uint64_t> tickets;
        std::vector<int32_t> sources;
        Buffer buf;
        int live = 0;  // queries not yet released
    };
    struct Lane {
        std::thread thr;
        std::shared_ptr<Batch> job;
        bool busy = false;
        bool long_job = false;  // the running batch was expected to take over long_batch_ms
        std::map<int, bool> has_workspace;
    };

    void scheduler_loop();
    void lane_loop(int lane);
    bool try_dispatch(std::unique_lock<std::mutex>& lk);
    int pick_graph_locked();
    int waiting_locked(int gid, int cap);
    void finish_batch(const std::shared_ptr<Batch>& b, int status, const std::string& err);
    void release_buffer_locked(Buffer buf);
    void fail_query_locked(Query& q, int status, const std::string& err);

    ServerConfig cfg_;
    std::unique_ptr<Executor> exec_;
    std::mutex mu_;
    std::condition_variable cv_sched_, cv_lane_, cv_done_;
    bool stop_ = false;
    std::vector<int> graphs_;
    std::unordered_map<uint64_t, Query> queries_;
    std::map<int, std::deque<uint64_t>> pending_;  // per graph, arrival order
    std::deque<uint64_t> arrival_order_;           // all pending tickets, arrival order
    std::vector<Lane> lanes_;
    std::map<int, std::vector<Buffer>> free_buffers_;  // pooled result buffers per graph
    int64_t used_ = 0;
    int64_t in_flight_ = 0;
    int64_t queued_ = 0;  // submitted, not yet dispatched, cancelled or failed
    uint64_t next_ticket_ = 1;
    // statistics
    int64_t submitted_ = 0, completed_ = 0, failed_ = 0, rejected_ = 0, batches_ = 0, batched_queries_ = 0;
    int64_t workspace_drops_ = 0;
    std::vector<double> latencies_ms_;
    Clock::time_point first_arrival_{}, last_done_{};
    bool have_first_ = false;
    std::thread sched_;
};

}  // namespace gg
