// Continuous-batching scheduler (see server.hpp).  Pure host C++: the device
// work happens behind Executor, so this file also builds into the host-only
// scheduler test.
#include "server.hpp"

#include <algorithm>
#include <limits>

namespace gg {

namespace {
double ms_between(Server::Clock::time_point a, Server::Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Timed waits use system_clock deadlines, the native clock of
// std::condition_variable in libstdc++ (pthread_cond_timedwait); relative
// waits go through pthread_cond_clockwait, which thread sanitizers of this
// toolchain do not model.
std::chrono::system_clock::time_point deadline(int timeout_ms) {
    return std::chrono::system_clock::now() + std::chrono::milliseconds(timeout_ms);
}
}  // namespace

Server::Server(const ServerConfig& cfg, std::unique_ptr<Executor> exec) : cfg_(cfg), exec_(std::move(exec)) {
    if (!exec_) fail(GB_EINVAL, "server needs an executor");
    if (cfg_.max_batch < 1 || cfg_.max_batch > 64) fail(GB_EINVAL, "max_batch must lie in [1, 64]");
    if (cfg_.lanes < 1 || cfg_.lanes > 16) fail(GB_EINVAL, "lanes must lie in [1, 16]");
    if (cfg_.queue_capacity < 1) fail(GB_EINVAL, "queue capacity must be positive");
    if (cfg_.plain) cfg_.max_batch = 1;
    if (cfg_.memory_budget <= 0) cfg_.memory_budget = std::numeric_limits<int64_t>::max();
    lanes_.resize(cfg_.lanes);
    for (int i = 0; i < cfg_.lanes; ++i) lanes_[i].thr = std::thread([this, i] { lane_loop(i); });
    sched_ = std::thread([this] { scheduler_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_sched_.notify_all();
    cv_lane_.notify_all();
    if (sched_.joinable()) sched_.join();
    for (auto& l : lanes_)
        if (l.thr.joinable()) l.thr.join();
    // Lanes finish their current batch before exiting; nothing is in flight.
    for (auto& kv : queries_) {
        Query& q = kv.second;
        if (q.batch && q.batch->buf.ptr) {
            exec_->free_result(q.batch->buf.ptr);
            q.batch->buf.ptr = nullptr;
        }
    }
    for (auto& kv : free_buffers_)
        for (auto& b : kv.second)
            if (b.ptr) exec_->free_result(b.ptr);
}

int Server::add_graph(int gid) {
    std::lock_guard<std::mutex> lk(mu_);
    if (gid < 0) gid = static_cast<int>(graphs_.size());
    if (std::find(graphs_.begin(), graphs_.end(), gid) != graphs_.end()) fail(GB_EINVAL, "graph already added");
    graphs_.push_back(gid);
    return gid;
}

uint64_t Server::submit(int gid, int32_t source) {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) fail(GB_EINVAL, "server is shutting down");
    if (std::find(graphs_.begin(), graphs_.end(), gid) == graphs_.end())
        fail(GB_EINVAL, "unknown graph id " + std::to_string(gid));
    const int64_t n = exec_->vertices(gid);
    if (source < 0 || source >= n)
        fail(GB_ERANGE, "source " + std::to_string(source) + " outside [0, " + std::to_string(n) + ")");
    if (queued_ >= cfg_.queue_capacity) {
        ++rejected_;
        fail(GB_EBUSY, "query queue is full");
    }
    const uint64_t t = next_ticket_++;
    Query& q = queries_[t];
    q.ticket = t;
    q.gid = gid;
    q.source = source;
    q.arrival = Clock::now();
    if (!have_first_) {
        first_arrival_ = q.arrival;
        have_first_ = true;
    }
    pending_[gid].push_back(t);
    arrival_order_.push_back(t);
    ++queued_;
    ++submitted_;
    cv_sched_.notify_one();
    return t;
}

void Server::scheduler_loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_sched_.wait(lk, [&] {
            if (stop_) return true;
            if (arrival_order_.empty()) return false;
            for (auto& l : lanes_)
                if (!l.busy) return true;
            return false;
        });
        if (stop_) break;
        if (!try_dispatch(lk)) cv_sched_.wait(lk);  // memory: wait for a completion or release
    }
    // Queries never dispatched fail at shutdown.
    for (uint64_t t : arrival_order_) {
        auto it = queries_.find(t);
        if (it != queries_.end() && it->second.status == GB_EAGAIN && !it->second.batch)
            fail_query_locked(it->second, GB_EINVAL, "server shut down before the query ran");
    }
    arrival_order_.clear();
    pending_.clear();
    cv_done_.notify_all();
}

void Server::fail_query_locked(Query& q, int status, const std::string& err) {
    q.status = status;
    q.error = err;
    q.done = Clock::now();
    ++failed_;
}

// Queries of graph gid waiting for a batch (stale entries at the front of
// its queue are dropped), counted up to `cap`.
int Server::waiting_locked(int gid, int cap) {
    std::deque<uint64_t>& pq = pending_[gid];
    auto waiting = [&](uint64_t t) {
        auto qi = queries_.find(t);
        return qi != queries_.end() && qi->second.status == GB_EAGAIN && !qi->second.batch;
    };
    while (!pq.empty() && !waiting(pq.front())) pq.pop_front();
    int k = 0;
    for (uint64_t t : pq)
        if (waiting(t) && ++k == cap) break;
    return k;
}

// The graph to start a batch on now, or -1.  Graphs whose batches are
// expected to run over long_batch_ms hold at most lanes - 1 lanes together,
// so a short query never waits behind them; when that cap is reached, their
// arrivals gather for the next batch.  Among the eligible graphs, the one
// whose oldest query has waited longest goes first.
int Server::pick_graph_locked() {
    int idle = 0, long_busy = 0;
    for (const Lane& l : lanes_) {
        if (!l.busy) ++idle;
        else if (l.long_job) ++long_busy;
    }
    if (idle == 0) return -1;
    int best = -1;
    Clock::time_point best_arrival{};
    for (auto& kv : pending_) {
        const int gid = kv.first;
        const int k = waiting_locked(gid, cfg_.max_batch);
        if (k == 0) continue;
        if (cfg_.lanes > 1 && long_busy >= cfg_.lanes - 1 && exec_->expected_ms(gid, k) > cfg_.long_batch_ms)
            continue;
        const Clock::time_point a = queries_[kv.second.front()].arrival;
        if (best < 0 || a < best_arrival) {
            best = gid;
            best_arrival = a;
        }
    }
    return best;
}

bool Server::try_dispatch(std::unique_lock<std::mutex>&) {
    while (!arrival_order_.empty()) {
        auto it = queries_.find(arrival_order_.front());
        if (it != queries_.end() && it->second.status == GB_EAGAIN && !it->second.batch) break;
        arrival_order_.pop_front();
    }
    if (arrival_order_.empty()) return true;
    const int gid = pick_graph_locked();
    if (gid < 0) return false;  // wait for a lane to finish (or a new arrival)
    std::deque<uint64_t>& pq = pending_[gid];

    int lane = -1;
    for (int i = 0; i < cfg_.lanes; ++i) {
        if (lanes_[i].busy) continue;
        if (lane < 0 || (lanes_[i].has_workspace.count(gid) && !lanes_[lane].has_workspace.count(gid))) lane = i;
    }
    if (lane < 0) return false;

    const int64_t n = exec_->vertices(gid);
    const int64_t row_bytes = n * int64_t(sizeof(int32_t));
    int k = waiting_locked(gid, cfg_.max_batch);  // queries of this graph that can go into the batch
    if (k == 0) return true;
    const int64_t ws_bytes =
        lanes_[lane].has_workspace.count(gid) ? 0 : static_cast<int64_t>(exec_->workspace_bytes(gid, cfg_.plain));
    std::vector<Buffer>& pool = free_buffers_[gid];
    auto pow2 = [](int x) {
        int p = 1;
        while (p < x) p <<= 1;
        return p;
    };
    // Smallest pooled buffer of this graph with at least `rows` rows.
    auto pooled = [&](int rows) {
        int best = -1;
        for (int i = 0; i < static_cast<int>(pool.size()); ++i)
            if (pool[i].rows >= rows && (best < 0 || pool[i].rows < pool[best].rows)) best = i;
        return best;
    };
    auto need_for = [&](int rows) { return ws_bytes + (pooled(rows) >= 0 ? 0 : pow2(rows) * row_bytes); };
    if (used_ + need_for(k) > cfg_.memory_budget) {
        // Give back pooled result buffers: other graphs' first, then this
        // graph's that are too small for the batch.
        for (auto& kv : free_buffers_) {
            if (kv.first == gid) continue;
            for (auto& b : kv.second) {
                exec_->free_result(b.ptr);
                used_ -= static_cast<int64_t>(b.bytes);
            }
            kv.second.clear();
        }
        for (auto it = pool.begin(); it != pool.end();) {
            if (it->rows < k) {
                exec_->free_result(it->ptr);
                used_ -= static_cast<int64_t>(it->bytes);
                it = pool.erase(it);
            } else {
                ++it;
            }
        }
        // Then the workspaces idle lanes hold for other graphs (an idle
        // lane's thread is waiting for a job and touches none of them).
        for (int i = 0; i < cfg_.lanes && used_ + need_for(k) > cfg_.memory_budget; ++i) {
            if (lanes_[i].busy) continue;
            for (auto it = lanes_[i].has_workspace.begin();
                 it != lanes_[i].has_workspace.end() && used_ + need_for(k) > cfg_.memory_budget;) {
                if (it->first == gid) {
                    ++it;
                    continue;
                }
                exec_->drop_workspace(i, it->first);
                used_ -= static_cast<int64_t>(exec_->workspace_bytes(it->first, cfg_.plain));
                ++workspace_drops_;
                it = lanes_[i].has_workspace.erase(it);
            }
        }
    }
    while (k > 1 && used_ + need_for(k) > cfg_.memory_budget) k = pow2(k) / 2;  // shrink to what fits
    const int64_t need = need_for(k);
    if (used_ + need > cfg_.memory_budget) {
        bool held = false;  // result buffers that a completion or release will return
        for (auto& kv : queries_)
            if (kv.second.batch && kv.second.status != GB_EAGAIN) held = true;
        if (in_flight_ == 0 && !held) {
            // Cannot fit even with everything else idle: fail this graph's queries.
            const std::string err = "a query on graph " + std::to_string(gid) + " needs " + std::to_string(need) +
                                    " bytes; the budget is " + std::to_string(cfg_.memory_budget) + " with " +
                                    std::to_string(used_) + " in use";
            for (uint64_t t : pq) {
                auto qi = queries_.find(t);
                if (qi != queries_.end() && qi->second.status == GB_EAGAIN && !qi->second.batch) {
                    fail_query_locked(qi->second, GB_ENOMEM, err);
                    --queued_;
                }
            }
            pq.clear();
            cv_done_.notify_all();
            return true;
        }
        return false;
    }

    auto b = std::make_shared<Batch>();
    b->gid = gid;
    b->lane = lane;
    while (!pq.empty() && static_cast<int>(b->tickets.size()) < k) {
        const uint64_t t = pq.front();
        pq.pop_front();
        auto qi = queries_.find(t);
        if (qi == queries_.end() || qi->second.status != GB_EAGAIN || qi->second.batch) continue;
        Query& q = qi->second;
        q.batch = b;
        --queued_;
        q.row = static_cast<int>(b->tickets.size());
        q.start = Clock::now();
        b->tickets.push_back(t);
        b->sources.push_back(q.source);
    }
    if (b->tickets.empty()) return true;
    used_ += need;
    if (ws_bytes) lanes_[lane].has_workspace[gid] = true;
    const int use = pooled(static_cast<int>(b->tickets.size()));
    if (use >= 0) {
        b->buf = pool[use];
        pool.erase(pool.begin() + use);
    } else {
        b->buf.rows = pow2(static_cast<int>(b->tickets.size()));
        b->buf.bytes = static_cast<size_t>(b->buf.rows * row_bytes);
        b->buf.gid = gid;
    }
    b->live = static_cast<int>(b->tickets.size());
    lanes_[lane].job = b;
    lanes_[lane].busy = true;
    lanes_[lane].long_job = exec_->expected_ms(gid, static_cast<int>(b->tickets.size())) > cfg_.long_batch_ms;
    ++in_flight_;
    ++batches_;
    batched_queries_ += static_cast<int64_t>(b->tickets.size());
    cv_lane_.notify_all();
    return true;
}

void Server::lane_loop(int lane) {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_lane_.wait(lk, [&] { return stop_ || lanes_[lane].job != nullptr; });
        std::shared_ptr<Batch> b = lanes_[lane].job;
        if (!b) break;  // stop with nothing to do
        lk.unlock();
        int status = GB_OK;
        std::string err;
        try {
            if (!b->buf.ptr) b->buf.ptr = exec_->alloc_result(b->buf.bytes);
            exec_->run(lane, b->gid, b->sources.data(), static_cast<int>(b->sources.size()),
                       static_cast<int32_t*>(b->buf.ptr), cfg_.plain);
        } catch (const Error& e) {
            status = e.code;
            err = e.what();
        } catch (const std::exception& e) {
            status = GB_EDEVICE;
            err = e.what();
        }
        lk.lock();
        finish_batch(b, status, err);
        lanes_[lane].job = nullptr;
        lanes_[lane].busy = false;
        lanes_[lane].long_job = false;
        --in_flight_;
        cv_sched_.notify_all();
        cv_done_.notify_all();
    }
}

void Server::finish_batch(const std::shared_ptr<Batch>& b, int status, const std::string& err) {
    const Clock::time_point now = Clock::now();
    last_done_ = now;
    int live = 0;
    for (uint64_t t : b->tickets) {
        auto qi = queries_.find(t);
        if (qi == queries_.end()) continue;
        Query& q = qi->second;
        q.done = now;
        if (status == GB_OK) {
            q.status = GB_OK;
            ++completed_;
            latencies_ms_.push_back(ms_between(q.arrival, now));
        } else {
            fail_query_locked(q, status, err);
        }
        if (q.released) {
            queries_.erase(qi);
        } else if (status == GB_OK) {
            ++live;
        }
    }
    b->live = live;
    if (status != GB_OK || live == 0) {
        if (!b->buf.ptr) {
            used_ -= static_cast<int64_t>(b->buf.bytes);  // never allocated
        } else {
            release_buffer_locked(b->buf);
        }
        b->buf.ptr = nullptr;
        for (uint64_t t : b->tickets) {  // failed queries hold no memory
            auto qi = queries_.find(t);
            if (qi != queries_.end()) qi->second.batch.reset();
        }
    }
}

void Server::release_buffer_locked(Buffer buf) {
    if (buf.ptr) free_buffers_[buf.gid].push_back(buf);
}

int Server::poll(uint64_t ticket, int32_t* host, QueryLatency* lat, const int32_t** device_row) {
    const int32_t* row = nullptr;
    int64_t n = 0;
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = queries_.find(ticket);
        if (it == queries_.end() || it->second.released) fail(GB_EINVAL, "unknown or released ticket");
        Query& q = it->second;
        if (q.status == GB_EAGAIN) return GB_EAGAIN;
        if (lat) {
            lat->queue_ms = q.batch || q.status == GB_OK ? ms_between(q.arrival, q.start) : 0;
            lat->run_ms = q.status == GB_OK ? ms_between(q.start, q.done) : 0;
            lat->total_ms = ms_between(q.arrival, q.done);
            lat->batch_size = q.batch ? static_cast<int>(q.batch->tickets.size()) : 0;
        }
        if (q.status != GB_OK) {
            set_last_error(q.error);
            return q.status;
        }
        n = exec_->vertices(q.gid);
        row = static_cast<const int32_t*>(q.batch->buf.ptr) + int64_t(q.row) * n;
    }
    if (device_row) *device_row = row;
    if (host) exec_->fetch(row, host, n);
    return GB_OK;
}

int Server::wait(uint64_t ticket, int timeout_ms, int32_t* host, QueryLatency* lat) {
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto ready = [&] {
            auto it = queries_.find(ticket);
            return it == queries_.end() || it->second.status != GB_EAGAIN;
        };
        if (timeout_ms < 0) {
            cv_done_.wait(lk, ready);
        } else if (!cv_done_.wait_until(lk, deadline(timeout_ms), ready)) {
            return GB_EAGAIN;
        }
    }
    return poll(ticket, host, lat, nullptr);
}

void Server::release(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = queries_.find(ticket);
    if (it == queries_.end()) return;
    Query& q = it->second;
    if (q.status == GB_EAGAIN) {
        if (!q.batch) {  // still queued: cancel
            auto& pq = pending_[q.gid];
            pq.erase(std::remove(pq.begin(), pq.end(), ticket), pq.end());
            queries_.erase(it);
            --queued_;
        } else {
            q.released = true;  // running: dropped when its batch finishes
        }
        return;
    }
    std::shared_ptr<Batch> b = q.batch;
    queries_.erase(it);
    if (b && b->buf.ptr && --b->live == 0) {
        release_buffer_locked(b->buf);
        b->buf.ptr = nullptr;
        cv_sched_.notify_all();
    }
}

bool Server::drain(int timeout_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto idle = [&] {
        if (in_flight_ > 0) return false;
        for (uint64_t t : arrival_order_) {
            auto it = queries_.find(t);
            if (it != queries_.end() && it->second.status == GB_EAGAIN) return false;
        }
        return true;
    };
    if (timeout_ms < 0) {
        cv_done_.wait(lk, idle);
        return true;
    }
    return cv_done_.wait_until(lk, deadline(timeout_ms), idle);
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ServerStats s;
    s.submitted = submitted_;
    s.completed = completed_;
    s.failed = failed_;
    s.rejected = rejected_;
    s.batches = batches_;
    s.mean_batch = batches_ ? double(batched_queries_) / double(batches_) : 0.0;
    if (!latencies_ms_.empty()) {
        std::vector<double> v(latencies_ms_);
        auto pct = [&](double p) {
            const size_t idx = std::min(v.size() - 1, static_cast<size_t>(p * double(v.size() - 1) + 0.5));
            std::nth_element(v.begin(), v.begin() + idx, v.end());
            return v[idx];
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        s.max_ms = *std::max_element(v.begin(), v.end());
        const double span = ms_between(first_arrival_, last_done_);
        s.qps = span > 0 ? double(completed_) / (span / 1e3) : 0.0;
    }
    s.memory_used = used_;
    s.workspace_drops = workspace_drops_;
    s.memory_budget = cfg_.memory_budget == std::numeric_limits<int64_t>::max() ? 0 : cfg_.memory_budget;
    return s;
}

void Server::reset_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    submitted_ = completed_ = failed_ = rejected_ = batches_ = batched_queries_ = workspace_drops_ = 0;
    latencies_ms_.clear();
    have_first_ = false;
}

const char* Server::query_error(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = queries_.find(ticket);
    if (it == queries_.end()) return "unknown ticket";
    set_last_error(it->second.error);
    return last_error();
}

}  // namespace gg
