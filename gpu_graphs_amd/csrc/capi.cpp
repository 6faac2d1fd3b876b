// extern "C" entry points declared in include/gpugraphs/gb.h.  Every call
// catches library exceptions, records the message for gb_last_error() and
// returns the status; nothing throws across the ABI and nothing exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>

#include "common.hpp"
#include "device_graph.hpp"
#include "handles.hpp"
#include "host_graph.hpp"

namespace gg {

template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return GB_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return GB_ENOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GB_EINVAL;
    } catch (...) {
        set_last_error("unknown error");
        return GB_EINVAL;
    }
}

void need(bool ok, const char* what) {
    if (!ok) fail(GB_EINVAL, what);
}

void check_device_ptr(int device, const void* p, const char* what) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(GB_EINVAL, std::string(what) + " is not a device pointer");
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)
        fail(GB_EINVAL, std::string(what) + " is not device memory");
    if (attr.device != device) fail(GB_EINVAL, std::string(what) + " lives on another device");
}

}  // namespace gg

using gg::guarded;
using gg::need;

extern "C" {

const char* gb_last_error(void) { return gg::last_error(); }
const char* gb_version(void) { return "gpugraphs-mi355x 0.1 (gfx950)"; }

int gb_graph_load_edgelist(const char* path, uint32_t flags, gb_graph** out) {
    return guarded([&] {
        need(path && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_graph>();
        h->g = gg::load_edgelist(path, flags);
        *out = h.release();
    });
}

int gb_graph_from_edges(int64_t n, int64_t m, const int32_t* src, const int32_t* dst, uint32_t flags,
                        gb_graph** out) {
    return guarded([&] {
        need(out && (m == 0 || (src && dst)), "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_graph>();
        h->g = gg::build_csr(n, src, dst, m, flags);
        *out = h.release();
    });
}

int gb_graph_from_csr(int64_t n, const int64_t* offsets, const int32_t* cols, uint32_t flags, gb_graph** out) {
    return guarded([&] {
        need(out && offsets, "null argument");
        *out = nullptr;
        need(n >= 0 && n <= INT32_MAX, "vertex count out of range");
        need(offsets[0] == 0, "offsets[0] must be 0");
        for (int64_t v = 0; v < n; ++v) need(offsets[v + 1] >= offsets[v], "offsets must be non-decreasing");
        const int64_t m = offsets[n];
        need(m == 0 || cols, "null cols");
        for (int64_t i = 0; i < m; ++i)
            if (cols[i] < 0 || cols[i] >= n) gg::fail(GB_ERANGE, "column id outside [0, n)");
        auto h = std::make_unique<gb_graph>();
        h->g.n = n;
        h->g.flags = flags;
        h->g.offsets.assign(offsets, offsets + n + 1);
        h->g.cols.assign(cols, cols + m);
        if (flags & GB_SYMMETRIZE) {
            // Rebuild from the edge list so that reverse edges are added.
            std::vector<int32_t> s(m), d(h->g.cols);
            for (int64_t v = 0; v < n; ++v)
                for (int64_t j = offsets[v]; j < offsets[v + 1]; ++j) s[j] = static_cast<int32_t>(v);
            h->g = gg::build_csr(n, s.data(), d.data(), m, flags);
        } else if (!(flags & GB_KEEP_ORDER)) {
            gg::sort_dedup_rows(h->g, flags & GB_DROP_SELF_LOOPS);
        }
        *out = h.release();
    });
}

int gb_graph_generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int permute_ids,
                           gb_graph** out) {
    return guarded([&] {
        need(out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_graph>();
        h->g = gg::generate_rmat(scale, edge_factor, a, b, c, seed, permute_ids != 0);
        *out = h.release();
    });
}

int gb_graph_generate_er(int64_t n_candidates, int64_t n_isolated, int64_t m_undirected, uint64_t seed,
                         gb_graph** out) {
    return guarded([&] {
        need(out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_graph>();
        h->g = gg::generate_er(n_candidates, n_isolated, m_undirected, seed);
        *out = h.release();
    });
}

int gb_graph_generate_grid(int64_t width, int64_t height, double extra_fraction, uint64_t seed, gb_graph** out) {
    return guarded([&] {
        need(out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_graph>();
        h->g = gg::generate_grid(width, height, extra_fraction, seed);
        *out = h.release();
    });
}

int gb_graph_info(const gb_graph* g, int64_t* n, int64_t* m) {
    return guarded([&] {
        need(g, "null graph");
        if (n) *n = g->g.n;
        if (m) *m = g->g.m();
    });
}

int gb_graph_csr(const gb_graph* g, const int64_t** offsets, const int32_t** cols) {
    return guarded([&] {
        need(g, "null graph");
        if (offsets) *offsets = g->g.offsets.data();
        if (cols) *cols = g->g.cols.data();
    });
}

int gb_graph_degree_buckets(gb_graph* g, int64_t* offsets, const int32_t** vertices) {
    return guarded([&] {
        need(g, "null graph");
        static_assert(GB_DEGREE_BUCKETS == gg::kDegreeBuckets, "bucket count");
        if (g->g.bucket_offsets.empty()) gg::compute_degree_buckets(g->g);
        if (offsets) std::copy(g->g.bucket_offsets.begin(), g->g.bucket_offsets.end(), offsets);
        if (vertices) *vertices = g->g.bucket_vertices.data();
    });
}

int gb_graph_is_symmetric(const gb_graph* g, int* symmetric) {
    return guarded([&] {
        need(g && symmetric, "null argument");
        *symmetric = gg::is_symmetric(g->g) ? 1 : 0;
    });
}

int gb_graph_write_edgelist(const gb_graph* g, const char* path) {
    return guarded([&] {
        need(g && path, "null argument");
        gg::write_edgelist(g->g, path);
    });
}

int gb_graph_save_binary(const gb_graph* g, const char* path) {
    return guarded([&] {
        need(g && path, "null argument");
        gg::save_binary(g->g, path);
    });
}

int gb_graph_load_binary(const char* path, gb_graph** out) {
    return guarded([&] {
        need(path && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_graph>();
        h->g = gg::load_binary(path);
        *out = h.release();
    });
}

void gb_graph_free(gb_graph* g) { delete g; }

int gb_device_count(int* count) {
    return guarded([&] {
        need(count, "null argument");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        if (e == hipErrorNoDevice) {
            (void)hipGetLastError();
            c = 0;
        } else if (e != hipSuccess) {
            (void)hipGetLastError();
            gg::fail(GB_EDEVICE, std::string("hipGetDeviceCount failed: ") + hipGetErrorString(e));
        }
        *count = c;
    });
}

int gb_device_upload(const gb_graph* g, int device, gb_dgraph** out) {
    return guarded([&] {
        need(g && out, "null argument");
        *out = nullptr;
        auto h = std::make_unique<gb_dgraph>();
        h->d = std::make_unique<gg::DeviceGraph>(g->g, device);
        *out = h.release();
    });
}

int gb_dgraph_info(const gb_dgraph* dg, int64_t* n, int64_t* m, int* device) {
    return guarded([&] {
        need(dg, "null device graph");
        if (n) *n = dg->d->n;
        if (m) *m = dg->d->m;
        if (device) *device = dg->d->device;
    });
}

int gb_dgraph_memory(const gb_dgraph* dg, int64_t* bytes) {
    return guarded([&] {
        need(dg && bytes, "null argument");
        *bytes = static_cast<int64_t>(dg->d->memory_bytes());
    });
}

void gb_device_free(gb_dgraph* dg) { delete dg; }

int gb_bfs_device(gb_dgraph* dg, int32_t source, int32_t* d_levels, int32_t* d_parents, gb_stats* st) {
    return guarded([&] {
        need(dg, "null device graph");
        need(d_levels, "null levels buffer");
        gg::check_device_ptr(dg->d->device, d_levels, "levels");
        if (d_parents) gg::check_device_ptr(dg->d->device, d_parents, "parents");
        dg->d->bfs(source, d_levels, d_parents, st);
    });
}

int gb_bfs(gb_dgraph* dg, int32_t source, int32_t* levels, int32_t* parents, gb_stats* st) {
    return guarded([&] {
        need(dg, "null device graph");
        need(levels, "null levels buffer");
        dg->d->bfs_to_host(source, levels, parents, st);
    });
}

int gb_bfs_config(gb_dgraph* dg, int block, int edges_per_thread, int blocks_per_cu, int mode, int64_t dense_edges) {
    return guarded([&] {
        need(dg, "null device graph");
        need(gg::bfs_config_supported(block, edges_per_thread), "unsupported block / edges_per_thread combination");
        need(blocks_per_cu >= 0, "blocks_per_cu must be >= 0");
        need(mode == GB_MODE_ASYNC || mode == GB_MODE_LEVEL_LOOP || mode == GB_MODE_GRAPH, "unknown mode");
        need(dense_edges >= 0, "dense_edges must be >= 0");
        gg::BfsConfig& c = dg->d->cfg;
        c.block = block;
        c.edges_per_thread = edges_per_thread;
        c.blocks_per_cu = blocks_per_cu;
        c.mode = mode;
        c.dense_edges = dense_edges > 0 ? dense_edges : dg->d->default_dense_edges();
    });
}

int gb_bfs_direction_config(gb_dgraph* dg, int enable, int alpha, int beta) {
    return guarded([&] {
        need(dg, "null device graph");
        need(alpha >= 0 && beta >= 0, "alpha and beta must be >= 0");
        gg::BfsConfig& c = dg->d->cfg;
        c.direction_opt = enable != 0;
        if (alpha) c.alpha = alpha;
        if (beta) c.beta = beta;
        dg->d->runner->reset_bottom_up_window();
    });
}

int gb_bfs_get_direction_config(const gb_dgraph* dg, int* enable, int* alpha, int* beta, int* symmetric) {
    return guarded([&] {
        need(dg, "null device graph");
        const gg::BfsConfig& c = dg->d->cfg;
        if (enable) *enable = c.direction_opt ? 1 : 0;
        if (alpha) *alpha = c.alpha;
        if (beta) *beta = c.beta;
        if (symmetric) *symmetric = dg->d->symmetric ? 1 : 0;
    });
}

int gb_bfs_last_bottom_up(const gb_dgraph* dg, int* first_level, int* last_level) {
    return guarded([&] {
        need(dg, "null device graph");
        if (first_level) *first_level = dg->d->runner->last_bu_first;
        if (last_level) *last_level = dg->d->runner->last_bu_last;
    });
}

int gb_bfs_launch_hints(const gb_dgraph* dg, int* depth, int* bu_lo, int* bu_hi) {
    return guarded([&] {
        need(dg, "null device graph");
        const gg::Runner& r = *dg->d->runner;
        if (depth) *depth = r.depth_hint;
        if (bu_lo) *bu_lo = r.bu_lo;
        if (bu_hi) *bu_hi = r.bu_hi;
    });
}

int gb_bfs_get_config(const gb_dgraph* dg, int* block, int* edges_per_thread, int* blocks_per_cu, int* mode,
                      int64_t* dense_edges) {
    return guarded([&] {
        need(dg, "null device graph");
        const gg::BfsConfig& c = dg->d->cfg;
        if (block) *block = c.block;
        if (edges_per_thread) *edges_per_thread = c.edges_per_thread;
        if (blocks_per_cu) *blocks_per_cu = c.blocks_per_cu;
        if (mode) *mode = c.mode;
        if (dense_edges) *dense_edges = c.dense_edges;
    });
}

int gb_bfs_batch(gb_dgraph* dg, const int32_t* sources, int k, int32_t* levels, gb_batch_stats* st) {
    return guarded([&] {
        need(dg && sources && levels, "null argument");
        dg->d->bfs_batch_to_host(sources, k, levels, st);
    });
}

int gb_bfs_batch_device(gb_dgraph* dg, const int32_t* sources, int k, int32_t* d_levels, gb_batch_stats* st) {
    return guarded([&] {
        need(dg && sources && d_levels, "null argument");
        gg::check_device_ptr(dg->d->device, d_levels, "levels");
        dg->d->bfs_batch(sources, k, d_levels, st);
    });
}

int gb_batch_config(gb_dgraph* dg, int block, int edges_per_thread, int use_graph, int64_t sparse_edges) {
    return guarded([&] {
        need(dg, "null device graph");
        need(gg::ms_config_supported(block, edges_per_thread), "unsupported block / edges_per_thread combination");
        need(sparse_edges >= 0, "sparse_edges must be >= 0");
        dg->d->ms_cfg.block = block;
        dg->d->ms_cfg.edges_per_thread = edges_per_thread;
        dg->d->ms_cfg.use_graph = use_graph != 0;
        dg->d->ms_cfg.sparse_edges = sparse_edges;
    });
}

int gb_batch_get_config(const gb_dgraph* dg, int* block, int* edges_per_thread, int* use_graph,
                        int64_t* sparse_edges) {
    return guarded([&] {
        need(dg, "null device graph");
        if (block) *block = dg->d->ms_cfg.block;
        if (edges_per_thread) *edges_per_thread = dg->d->ms_cfg.edges_per_thread;
        if (use_graph) *use_graph = dg->d->ms_cfg.use_graph ? 1 : 0;
        if (sparse_edges)
            *sparse_edges = dg->d->ms_cfg.sparse_edges > 0 ? dg->d->ms_cfg.sparse_edges
                                                           : dg->d->default_ms_sparse_edges();
    });
}

int gb_batch_direction_config(gb_dgraph* dg, int enable, int alpha, int beta) {
    return guarded([&] {
        need(dg, "null device graph");
        need(alpha >= 0 && beta >= -1, "alpha must be >= 0 and beta >= -1");
        dg->d->ms_cfg.direction_opt = enable != 0;
        if (alpha) dg->d->ms_cfg.alpha = alpha;
        if (beta >= 0) dg->d->ms_cfg.beta = beta;
    });
}

int gb_batch_get_direction_config(const gb_dgraph* dg, int* enable, int* alpha, int* beta) {
    return guarded([&] {
        need(dg, "null device graph");
        if (enable) *enable = dg->d->ms_cfg.direction_opt ? 1 : 0;
        if (alpha) *alpha = dg->d->ms_cfg.alpha;
        if (beta) *beta = dg->d->ms_cfg.beta;
    });
}

int gb_batch_last_bottom_up(const gb_dgraph* dg, int* levels) {
    return guarded([&] {
        need(dg && levels, "null argument");
        *levels = dg->d->runner->ms.last_bu_levels;
    });
}

int gb_batch_deep_config(gb_dgraph* dg, int mode) {
    return guarded([&] {
        need(dg, "null device graph");
        need(mode == GB_DEEP_AUTO || mode == GB_DEEP_ON || mode == GB_DEEP_OFF, "unknown deep mode");
        dg->d->ms_cfg.deep = mode;
    });
}

int gb_batch_get_deep_config(const gb_dgraph* dg, int* mode, int* used) {
    return guarded([&] {
        need(dg, "null device graph");
        if (mode) *mode = dg->d->ms_cfg.deep;
        if (used) *used = dg->d->runner->ms.last_deep ? 1 : 0;
    });
}

int gb_bfs_deep_config(gb_dgraph* dg, int mode) {
    return guarded([&] {
        need(dg, "null device graph");
        need(mode == GB_DEEP_AUTO || mode == GB_DEEP_ON || mode == GB_DEEP_OFF, "unknown deep mode");
        dg->d->cfg.deep = mode;
    });
}

int gb_bfs_get_deep_config(const gb_dgraph* dg, int* mode, int* used) {
    return guarded([&] {
        need(dg, "null device graph");
        if (mode) *mode = dg->d->cfg.deep;
        if (used) *used = dg->d->runner->last_deep ? 1 : 0;
    });
}

int gb_bfs_tail_config(gb_dgraph* dg, int level) {
    return guarded([&] {
        need(dg, "null device graph");
        need(level == GB_TAIL_AUTO || level == GB_TAIL_OFF || level >= 0, "unknown tail level");
        dg->d->cfg.tail = level;
    });
}

int gb_bfs_get_tail_config(const gb_dgraph* dg, int* level, int* used, int* outcome) {
    return guarded([&] {
        need(dg, "null device graph");
        if (level) *level = dg->d->cfg.tail;
        if (used) *used = dg->d->runner->last_tail_at;
        if (outcome) *outcome = dg->d->runner->last_tail_outcome;
    });
}

int gb_bfs_autotune(gb_dgraph* dg, const int32_t* sources, int count, int reps, gb_tune_result* out) {
    return guarded([&] {
        need(dg && sources, "null argument");
        gg::autotune_bfs(*dg->d, sources, count, reps, out);
    });
}

int gb_batch_autotune(gb_dgraph* dg, const int32_t* sources, int k, int reps, gb_tune_result* out) {
    return guarded([&] {
        need(dg && sources, "null argument");
        gg::autotune_batch(*dg->d, sources, k, reps, out);
    });
}

int gb_bfs_apply_tuning(gb_dgraph* dg, const gb_tune_result* t) {
    return guarded([&] {
        need(dg && t, "null argument");
        need(gg::bfs_config_supported(t->block, t->edges_per_thread), "unsupported block / edges_per_thread combination");
        need(t->alpha > 0 && t->beta >= 0 && t->frontier_edges > 0, "thresholds must be positive");
        gg::BfsConfig& c = dg->d->cfg;
        c.block = t->block;
        c.edges_per_thread = t->edges_per_thread;
        c.alpha = t->alpha;
        if (t->beta) c.beta = t->beta;
        c.dense_edges = t->frontier_edges;
        dg->d->runner->reset_bottom_up_window();
    });
}

int gb_batch_apply_tuning(gb_dgraph* dg, const gb_tune_result* t) {
    return guarded([&] {
        need(dg && t, "null argument");
        need(gg::ms_config_supported(t->block, t->edges_per_thread), "unsupported block / edges_per_thread combination");
        need(t->alpha > 0 && t->low_degree > 0 && t->frontier_edges > 0, "thresholds must be positive");
        gg::MsConfig& c = dg->d->ms_cfg;
        c.block = t->block;
        c.edges_per_thread = t->edges_per_thread;
        c.alpha = t->alpha;
        c.beta = t->beta;
        c.low_degree = t->low_degree;
        c.sparse_edges = t->frontier_edges;
    });
}

int gb_exclusive_scan_i32(int device, const int32_t* d_in, int32_t* d_out, int64_t n) {
    return guarded([&] {
        if (n > 0) {
            gg::check_device_ptr(device, d_in, "scan input");
            gg::check_device_ptr(device, d_out, "scan output");
        }
        gg::exclusive_scan_i32(device, d_in, d_out, n);
    });
}

int gb_exclusive_scan_i32_noalloc(int device, const int32_t* d_in, int32_t* d_out, int64_t n) {
    return guarded([&] {
        if (n > 0) {
            gg::check_device_ptr(device, d_in, "scan input");
            gg::check_device_ptr(device, d_out, "scan output");
        }
        gg::exclusive_scan_i32_noalloc(device, d_in, d_out, n);
    });
}

int64_t gb_exclusive_scan_noalloc_max(void) { return gg::exclusive_scan_noalloc_max(); }

int gb_dmalloc(int device, size_t bytes, void** ptr) {
    return guarded([&] {
        need(ptr, "null argument");
        *ptr = nullptr;
        gg::DeviceGuard guard(device);
        GG_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    });
}

int gb_dfree(int device, void* ptr) {
    return guarded([&] {
        gg::DeviceGuard guard(device);
        if (ptr) GG_HIP(hipFree(ptr));
    });
}

int gb_memcpy_htod(int device, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        gg::DeviceGuard guard(device);
        if (bytes) GG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

int gb_memcpy_dtoh(int device, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        gg::DeviceGuard guard(device);
        if (bytes) GG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

int gb_device_synchronize(int device) {
    return guarded([&] {
        gg::DeviceGuard guard(device);
        GG_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"
