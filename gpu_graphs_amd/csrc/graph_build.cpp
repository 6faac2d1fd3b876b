// Edge list -> CSR construction (count, scan, scatter, per-row sort/dedup).
//
// The reference builds `vector<vector<int>> adjList(n+1)`, copies each row
// (`std::vector<int> cur_node = adjList[i]`) and backfills empty rows
// (1D Partition/src/csr.cpp:25-65).  Here the CSR is built in three parallel
// passes over flat arrays; GB_KEEP_ORDER keeps the reference's row contents
// (file order, duplicates, self-loops) with a stable sequential scatter.
#include <algorithm>
#include <atomic>
#include <cstring>

#include "common.hpp"
#include "host_graph.hpp"

namespace gg {

namespace {

// Splits rows [0,n) into `parts` slices with about equal numbers of CSR
// entries, so hub rows of power-law graphs do not serialise one worker.
std::vector<int64_t> balanced_row_splits(const std::vector<int64_t>& offsets, int64_t n, int parts) {
    std::vector<int64_t> cut(parts + 1, n);
    cut[0] = 0;
    const int64_t m = offsets[n];
    for (int p = 1; p < parts; ++p) {
        int64_t target = m * p / parts;
        int64_t r = std::upper_bound(offsets.begin(), offsets.begin() + n + 1, target) - offsets.begin() - 1;
        r = std::max<int64_t>(r, cut[p - 1]);
        cut[p] = std::min<int64_t>(r, n);
    }
    return cut;
}

template <class F>
void parallel_rows(const std::vector<int64_t>& offsets, int64_t n, F&& fn) {
    int workers = host_threads();
    if (n < 4096 || workers == 1) {
        fn(int64_t(0), n);
        return;
    }
    auto cut = balanced_row_splits(offsets, n, workers);
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) {
        int64_t b = cut[w], e = cut[w + 1];
        if (b < e) pool.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto& t : pool) t.join();
}

void exclusive_scan_inplace(std::vector<int64_t>& a) {
    int64_t run = 0;
    for (auto& x : a) {
        int64_t v = x;
        x = run;
        run += v;
    }
}

}  // namespace

void sort_dedup_rows(HostGraph& g, bool drop_self_loops) {
    const int64_t n = g.n;
    std::vector<int64_t> newdeg(n + 1, 0);
    parallel_rows(g.offsets, n, [&](int64_t b, int64_t e) {
        for (int64_t v = b; v < e; ++v) {
            int32_t* lo = g.cols.data() + g.offsets[v];
            int32_t* hi = g.cols.data() + g.offsets[v + 1];
            std::sort(lo, hi);
            int32_t* out = lo;
            for (int32_t* p = lo; p < hi; ++p) {
                if (drop_self_loops && *p == v) continue;
                if (out != lo && out[-1] == *p) continue;
                *out++ = *p;
            }
            newdeg[v] = out - lo;
        }
    });
    std::vector<int64_t> noff(newdeg);
    exclusive_scan_inplace(noff);  // noff[n] = total because newdeg[n] = 0
    if (noff[n] == g.m()) return;  // nothing removed
    std::vector<int32_t> ncols(noff[n]);
    parallel_rows(noff, n, [&](int64_t b, int64_t e) {
        for (int64_t v = b; v < e; ++v)
            std::memcpy(ncols.data() + noff[v], g.cols.data() + g.offsets[v], sizeof(int32_t) * newdeg[v]);
    });
    g.offsets.swap(noff);
    g.cols.swap(ncols);
}

void compute_degree_buckets(HostGraph& g) {
    const int64_t n = g.n;
    const int workers = host_threads();
    // Per-worker counts over the contiguous vertex slices of parallel_for
    // (the same slices in both passes), then a scatter that keeps ids
    // ascending inside each bucket.
    std::vector<std::vector<int64_t>> cnt(workers, std::vector<int64_t>(kDegreeBuckets, 0));
    parallel_for(n, [&](int64_t b, int64_t e, int w) {
        for (int64_t v = b; v < e; ++v) cnt[w][degree_bucket(g.degree(v))]++;
    }, workers);
    g.bucket_offsets.assign(kDegreeBuckets + 1, 0);
    std::vector<std::vector<int64_t>> start(workers, std::vector<int64_t>(kDegreeBuckets, 0));
    int64_t pos = 0;
    for (int b = 0; b < kDegreeBuckets; ++b) {
        g.bucket_offsets[b] = pos;
        for (int w = 0; w < workers; ++w) {
            start[w][b] = pos;
            pos += cnt[w][b];
        }
    }
    g.bucket_offsets[kDegreeBuckets] = pos;
    g.bucket_vertices.resize(n);
    parallel_for(n, [&](int64_t b, int64_t e, int w) {
        std::vector<int64_t> cur = start[w];
        for (int64_t v = b; v < e; ++v) g.bucket_vertices[cur[degree_bucket(g.degree(v))]++] = static_cast<int32_t>(v);
    }, workers);
}

bool is_symmetric(const HostGraph& g) {
    if (g.symmetric) return true;
    if (g.flags & GB_KEEP_ORDER) return false;  // rows not sorted: not checked
    std::atomic<bool> ok{true};
    parallel_for(g.n, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; ++v) {
            if (!ok.load(std::memory_order_relaxed)) return;
            for (int64_t i = g.offsets[v]; i < g.offsets[v + 1]; ++i) {
                const int32_t u = g.cols[i];
                const int32_t* rb = g.cols.data() + g.offsets[u];
                const int32_t* re = g.cols.data() + g.offsets[u + 1];
                if (!std::binary_search(rb, re, static_cast<int32_t>(v))) {
                    ok.store(false, std::memory_order_relaxed);
                    return;
                }
            }
        }
    });
    return ok.load();
}

HostGraph build_csr(int64_t n, const int32_t* src, const int32_t* dst, int64_t m, uint32_t flags) {
    if (n < 0 || n > INT32_MAX) fail(GB_EINVAL, "vertex count must lie in [0, 2^31)");
    if (m < 0) fail(GB_EINVAL, "negative edge count");
    const bool sym = flags & GB_SYMMETRIZE;
    const bool keep = flags & GB_KEEP_ORDER;
    const bool drop_loops = flags & GB_DROP_SELF_LOOPS;

    // Bounds check (the reference indexes adjList[from] unchecked, csr.cpp:28).
    {
        std::atomic<int64_t> bad{-1};
        parallel_for(m, [&](int64_t b, int64_t e, int) {
            for (int64_t i = b; i < e; ++i)
                if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) {
                    bad.store(i);
                    return;
                }
        });
        if (bad.load() >= 0) {
            int64_t i = bad.load();
            fail(GB_ERANGE, "edge " + std::to_string(i) + " (" + std::to_string(src[i]) + " -> " +
                                std::to_string(dst[i]) + ") has an id outside [0, " + std::to_string(n) + ")");
        }
    }

    HostGraph g;
    g.n = n;
    g.flags = flags;
    g.symmetric = sym;
    g.offsets.assign(n + 1, 0);

    if (keep) {
        // Stable counting sort by source: row v lists its edges in input order,
        // duplicates and self-loops kept (unless asked to drop self-loops).
        for (int64_t i = 0; i < m; ++i) {
            if (drop_loops && src[i] == dst[i]) continue;
            g.offsets[src[i]]++;
            if (sym) g.offsets[dst[i]]++;
        }
        exclusive_scan_inplace(g.offsets);
        g.cols.resize(g.offsets[n]);
        std::vector<int64_t> cur(g.offsets.begin(), g.offsets.end() - 1);
        for (int64_t i = 0; i < m; ++i) {
            if (drop_loops && src[i] == dst[i]) continue;
            g.cols[cur[src[i]]++] = dst[i];
            if (sym) g.cols[cur[dst[i]]++] = src[i];
        }
        return g;
    }

    // Parallel count with relaxed atomic increments.
    int64_t* cnt = g.offsets.data();
    parallel_for(m, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            if (drop_loops && src[i] == dst[i]) continue;
            __atomic_fetch_add(&cnt[src[i]], 1, __ATOMIC_RELAXED);
            if (sym) __atomic_fetch_add(&cnt[dst[i]], 1, __ATOMIC_RELAXED);
        }
    });
    exclusive_scan_inplace(g.offsets);
    g.cols.resize(g.offsets[n]);
    std::vector<int64_t> cur(g.offsets.begin(), g.offsets.end() - 1);
    int64_t* c = cur.data();
    int32_t* cols = g.cols.data();
    parallel_for(m, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            if (drop_loops && src[i] == dst[i]) continue;
            cols[__atomic_fetch_add(&c[src[i]], 1, __ATOMIC_RELAXED)] = dst[i];
            if (sym) cols[__atomic_fetch_add(&c[dst[i]], 1, __ATOMIC_RELAXED)] = src[i];
        }
    });
    cur.clear();
    cur.shrink_to_fit();
    sort_dedup_rows(g, drop_loops);
    return g;
}

}  // namespace gg
