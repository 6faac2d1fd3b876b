// Device-resident graph and the single-source BFS driver.
//
// Replaces gpuBFS (1D Partition/include/gpuBFS.cuh:9-54, gpuBFS.cu:13-208).
// The reference splits vertices over NVSHMEM PEs and copies a CSR slice to
// each (init_device, gpuBFS.cu:139-185); here the whole graph lives in the HBM
// of one GPU and the device buffers are allocated once per handle, not per
// BFS.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_graph.hpp"

namespace gg {

#define GG_HIP(call)                                                                                        \
    do {                                                                                                    \
        hipError_t _e = (call);                                                                             \
        if (_e != hipSuccess)                                                                               \
            ::gg::fail(_e == hipErrorOutOfMemory ? GB_ENOMEM : GB_EDEVICE,                                  \
                       std::string(#call) + " failed: " + hipGetErrorString(_e) + " (" + __FILE__ + ":" + \
                           std::to_string(__LINE__) + ")");                                                 \
    } while (0)

// Makes `device` current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int device) {
        GG_HIP(hipGetDevice(&prev));
        if (prev != device) GG_HIP(hipSetDevice(device));
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Tunables of the single-source top-down kernels.
struct BfsConfig {
    int block = 256;              // threads per workgroup of expand() (4 waves)
    int edges_per_thread = 4;     // K: a tile is block * K frontier edges
    int blocks_per_cu = 0;        // 0 = at most 16 waves per CU
    int mode = GB_MODE_ASYNC;     // direct launches: cheaper than a replayed graph for ~10 kernels
    int64_t dense_edges = 0;      // levels with more frontier edges use bitmap compaction
    bool direction_opt = true;    // bottom-up levels on symmetric graphs
    int alpha = 14;               // bottom-up when the frontier grows and its edges * alpha > unexplored edges
    int beta = 24;                // back top-down when the frontier shrinks below n / beta vertices
    int deep = GB_DEEP_AUTO;      // one launch per level for deep searches with thin frontiers
    int tail = GB_TAIL_AUTO;      // level from which one workgroup finishes the search (GB_TAIL_*, or >= 0)
};

// Per-call parameters of a single-source BFS, read by the kernels from
// device memory so that one captured launch sequence serves every call.
struct BfsParams {
    int32_t source;
    int32_t flags;     // BFS_PARAM_* (direct launches only; a replayed graph reads 0)
    int32_t* levels;
    int32_t* parents;  // may be null
};

// Device workspace of one BFS (reused across calls on the same handle).
struct BfsWorkspace {
    uint32_t* vis = nullptr;                    // ceil(n/32) words: claimed vertices
    uint32_t* prev = nullptr;                   // ceil(n/32) words: vertices already queued
    uint32_t* fb[3] = {nullptr, nullptr, nullptr};  // ceil(n/32) words each: level L's frontier bitmap at fb[L % 3]
    int64_t* q_rb[2] = {nullptr, nullptr};      // frontier: first CSR entry of the vertex
    uint32_t* q_eoff[2] = {nullptr, nullptr};   // frontier: exclusive prefix of degrees in the level
    int32_t* q_v[2] = {nullptr, nullptr};       // frontier: vertex id
    uint32_t* tile_start[2] = {nullptr, nullptr};  // frontier entry that owns edge t*T
    unsigned long long* ctl = nullptr;          // counters ring, stats, flags (see bfs.hip)
    int32_t* levels = nullptr;                  // levels staging for host-pointer calls
    int32_t* parents = nullptr;
    BfsParams* params = nullptr;                // device
    BfsParams* h_params = nullptr;              // pinned, coherent: written by the host, read by the clear kernel
    BfsParams* h_params_dev = nullptr;          // its device-side address
    int64_t tile_cap = 0;                       // entries per tile_start buffer
    int64_t nchunks = 0;                        // compaction chunks
    int tile_edges = 0;                         // T the tile_start buffers were sized for
    size_t bytes = 0;
    // Captured launch sequence (GB_MODE_GRAPH) and what it was captured for.
    hipGraphExec_t graph_exec = nullptr;
    int g_depth = 0, g_lo = 0, g_hi = 0, g_block = 0, g_k = 0, g_bu = 0, g_alpha = 0, g_beta = 0, g_deep = 0,
        g_tail = -1;
    int64_t g_dense = -1;
};

// Tunables of the batched multi-source kernels.
struct MsConfig {
    int block = 256;
    int edges_per_thread = 2;
    bool use_graph = true;      // replay one captured hipGraph per batch
    int64_t sparse_edges = 0;   // levels with at most this many frontier edges pass a touched list
    bool direction_opt = true;  // bottom-up levels on symmetric graphs
    int alpha = 2;              // bottom-up when frontier edges * alpha > edges not yet fully seen
    int beta = 0;               // ... and the frontier holds >= (vertices with edges) / beta vertices (0: any)
    int low_degree = 32;        // bottom-up: rows up to this long are scanned by one lane
    int deep = GB_DEEP_AUTO;    // one fused launch per level for deep batches with thin frontiers
};

// Per-batch parameters, read by the kernels from device memory so that one
// captured sequence serves every batch (any k, any output buffer).
struct MsParams {
    int32_t sources[64];
    int32_t k;
    int32_t clear_all;              // clear the state of every vertex (else of those with edges)
    int32_t* levels;                // k rows of n int32
    int32_t prev_sources[64];       // the previous batch's sources (their state is cleared)
    int32_t prev_k;
    int32_t bu_lo;                  // levels [bu_lo, bu_hi] go bottom-up, and their update builds no queue
    int32_t bu_hi;
    int32_t pad;
};

// Device workspace of the batched multi-source BFS.
struct MsWorkspace {
    unsigned long long* sn = nullptr;     // 2n: per vertex, sources that reached it and sources reaching it now
    unsigned long long* fmask = nullptr;  // n: copy of seen, stored when it changes (gathered bottom-up)
    unsigned long long* fmask2 = nullptr; // n: the other copy (bottom-up levels that update inline write it);
                                          // next1's memory, not freed on its own
    unsigned long long* lv8 = nullptr;    // 8n: one level byte per (vertex, source)
    unsigned long long* reached = nullptr;  // ceil(n/64): vertices whose lv8 line the batch has written
    int64_t* q_rb = nullptr;              // queue of active vertices
    uint32_t* q_eoff = nullptr;
    int32_t* q_v = nullptr;
    unsigned long long* q_fm = nullptr;   // their frontier masks
    uint32_t* tile_start = nullptr;
    int32_t* touched = nullptr;           // n: vertices first reached in a sparse level
    int32_t* touched2 = nullptr;          // n: deep batches: the list of odd levels (touched: even)
    unsigned long long* next1 = nullptr;  // n: deep batches: next of odd levels (even: sn's next)
    unsigned long long* ctl = nullptr;
    MsParams* params = nullptr;           // device
    MsParams* h_params = nullptr;         // pinned, coherent: written by the host, read by the clear kernel
    MsParams* h_params_dev = nullptr;     // its device-side address
    unsigned long long* h_ctl = nullptr;  // pinned, coherent: 2 snapshots of ctl
    unsigned long long* h_ctl_dev = nullptr;  // its device-side address (the kernels store the snapshots)
    unsigned long long* partial = nullptr;  // per-workgroup statistics
    int32_t* levels = nullptr;            // staging for host-pointer calls
    size_t levels_bytes = 0;
    int64_t tile_cap = 0;
    int64_t nchunks = 0;
    int tile_edges = 0;
    int last_depth = 0;
    int depth_hint = 0;                   // levels queued before the first done check (see bfs_batch)
    int short_streak = 0;                 // batches in a row shallower than depth_hint
    int last_bu_levels = 0;               // levels the last batch ran bottom-up
    int bu_lo = 1, bu_hi = 0;             // first / last level the last batch's direction rule chose bottom-up
    int bu_alpha = -1, bu_beta = -1;      // the thresholds that batch ran with
    bool tail_clean = false;              // vertices without edges hold no state but prev_sources'
    bool deep_learned = false;            // the last batch was deep with thin frontiers: fused levels next
    bool last_deep = false;               // the last batch ran fused levels
    int32_t prev_sources[64] = {};        // sources of the last batch that completed
    int prev_k = 0;
    size_t bytes = 0;
    hipGraphExec_t graph_exec = nullptr;  // captured batch sequence
    int graph_depth = 0, graph_block = 0, graph_kpt = 0, graph_bu = -1, graph_alpha = 0, graph_low = 0,
        graph_deep = -1, graph_beta = -1, graph_nib = -1;
    int64_t graph_sparse = -1;
};

struct DeviceGraph;

// Execution context of BFS calls on one graph: a HIP stream, its events and
// the single- and multi-source workspaces.  A DeviceGraph owns one, used by
// the gb_bfs* calls; the query server creates one per lane so that batches on
// the same graph can run concurrently on different streams.
struct Runner {
    DeviceGraph& g;
    const int device;
    const int64_t n, m;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // created (and destroyed) by the runner, else the caller's
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_kernel = nullptr;
    hipEvent_t ev_batch[2] = {nullptr, nullptr};
    static constexpr int kCtlWords = 32;
    static constexpr int kDefaultBuHigh = 15;
    unsigned long long* h_ctl = nullptr;  // pinned, coherent: 2 snapshots of the control block
    unsigned long long* h_ctl_dev = nullptr;  // its device-side address (the kernels store the snapshots)
    int last_num_levels = 0;              // depth of the previous BFS: first batch size
    int bu_lo = 1, bu_hi = kDefaultBuHigh;  // levels that get a bottom-up launch
    bool bu_learned = false;                // window learned from a search (else the default)
    int thin_streak = 0;                    // searches in a row without dense levels
    int last_bu_first = -1, last_bu_last = -1;  // levels the previous BFS searched bottom-up
    int depth_hint = 0;                     // levels launched before the first done check
    int short_streak = 0;                   // searches in a row shallower than depth_hint
    bool last_deep = false;                 // the last search ran in deep mode
    bool last_fused = false;                // the last search ran one launch per level
    int tail_hint = -1;                     // level from which one workgroup finishes the search (-1: none yet)
    static constexpr int kTailWindow = 64;  // searches the tail level is chosen over
    int tail_seen[kTailWindow] = {};        // where their final runs of small levels began (ring)
    int tail_seen_n = 0;                    // searches recorded so far
    int last_tail_at = -1;                  // level the last search started its tail at (-1: none)
    int last_tail_outcome = 0;              // GB_TAIL_* outcome of that tail (0: none)
    const void* occ_kernel = nullptr;       // expand kernel whose occupancy occ_blocks holds
    int occ_blocks = 0;
    BfsWorkspace ws;
    MsWorkspace ms;

    explicit Runner(DeviceGraph& graph);
    // A runner that launches on the caller's stream (which must outlive it):
    // the query server gives all runners of one lane the lane's stream.
    Runner(DeviceGraph& graph, hipStream_t lane_stream);
    ~Runner();
    Runner(const Runner&) = delete;
    Runner& operator=(const Runner&) = delete;

    // levels / parents are device pointers (parents may be null).
    void bfs(int32_t source, int32_t* d_levels, int32_t* d_parents, gb_stats* st);
    // Host-pointer variant: runs into staging buffers, then copies levels
    // (and parents) to the host.
    void bfs_to_host(int32_t source, int32_t* h_levels, int32_t* h_parents, gb_stats* st);
    // Batched multi-source BFS of k <= 64 sources (host array); levels is a
    // device buffer of k rows of n int32.
    void bfs_batch(const int32_t* h_sources, int k, int32_t* d_levels, gb_batch_stats* st);
    double last_ms = 0;  // device time of the last bfs / bfs_batch call
    double wall_khz = 0; // rate of the device clock (wall_clock64) in kHz
    void bfs_batch_to_host(const int32_t* h_sources, int k, int32_t* h_levels, gb_batch_stats* st);
    // Allocates the batch workspace now (the server charges it to its budget).
    void prepare_batch();

    size_t bytes() const { return ws.bytes + ms.bytes + ms.levels_bytes; }
    // Forgets the bottom-up launch window learned so far (after the direction
    // thresholds change).
    void reset_bottom_up_window() {
        bu_lo = 1;
        bu_hi = kDefaultBuHigh;
        bu_learned = false;
        thin_streak = 0;
    }

   private:
    // Frees everything the runner holds (destructor, failed constructor).
    void release();
    void learn_from_search(bool had_dense_level, unsigned long long tail_from);
    void ensure_workspace();
    void free_workspace();
    void ensure_ms_workspace();
    void free_ms_workspace();
};

struct DeviceGraph {
    int device = 0;
    int64_t n = 0, m = 0;
    int64_t n_edged = 0;  // vertices with at least one CSR entry
    std::vector<uint8_t> deg8;  // host: each vertex's degree, clipped at 255
    bool rows_unique = true;    // no row repeats a neighbour (built without GB_KEEP_ORDER)
    int64_t* offsets = nullptr;  // n + 1
    int32_t* cols = nullptr;     // m
    uint32_t* iso = nullptr;     // ceil(n/32) words: degree-0 vertices (symmetric graphs only)
    int4* head = nullptr;        // n x 4: first neighbours of each vertex, -1 padded (symmetric graphs only)
    // Degree-ordered copy for the batched search: vertex ids by descending
    // log2-degree bucket, so the per-vertex masks of the hubs that most rows
    // point to share cache lines.
    int64_t* ms_offsets = nullptr;  // n + 1
    int32_t* ms_cols = nullptr;     // m, in new ids (rows keep their entry order)
    int32_t* ms_perm = nullptr;     // n: original id -> new id
    int32_t* ms_iperm = nullptr;    // n: new id -> original id
    int64_t ms_heavy = 0;           // vertices [0, ms_heavy) have rows of >= kMsHeavyRow entries, by descending degree
    int64_t ms_nonzero = 0;         // symmetric graphs: vertices [ms_nonzero, n) have no edges (else n)
    int32_t ms_heavy_rounds = 0;    // tiles of kMsHeavyRow entries in the longest row
    uint32_t* ms_heavy_tiles = nullptr;  // ms_heavy_rounds + 1: first tile of each round (tile r of every row that has one)
    bool symmetric = false;      // every edge has its reverse: bottom-up search is valid
    size_t graph_bytes = 0;
    int num_cus = 0;
    BfsConfig cfg;
    MsConfig ms_cfg;
    std::unique_ptr<Runner> runner;  // context of the gb_bfs* calls

    DeviceGraph(const HostGraph& g, int device);
    ~DeviceGraph();
    DeviceGraph(const DeviceGraph&) = delete;
    DeviceGraph& operator=(const DeviceGraph&) = delete;

    void bfs(int32_t source, int32_t* d_levels, int32_t* d_parents, gb_stats* st) {
        runner->bfs(source, d_levels, d_parents, st);
    }
    void bfs_to_host(int32_t source, int32_t* h_levels, int32_t* h_parents, gb_stats* st) {
        runner->bfs_to_host(source, h_levels, h_parents, st);
    }
    void bfs_batch(const int32_t* h_sources, int k, int32_t* d_levels, gb_batch_stats* st) {
        runner->bfs_batch(h_sources, k, d_levels, st);
    }
    void bfs_batch_to_host(const int32_t* h_sources, int k, int32_t* h_levels, gb_batch_stats* st) {
        runner->bfs_batch_to_host(h_sources, k, h_levels, st);
    }

    size_t memory_bytes() const { return graph_bytes + (runner ? runner->bytes() : 0); }
    // Frees the device copies and the runner (destructor, failed constructor).
    void release();
    int64_t default_dense_edges() const { return std::max<int64_t>(int64_t(1) << 19, m / 256); }
    int64_t default_ms_sparse_edges() const { return std::max<int64_t>(int64_t(1) << 16, n / 16); }
};

bool bfs_config_supported(int block, int k);
// Rows this long or longer are "heavy" in the batched bottom-up step: their
// entries are scanned in tiles of this size spread over all workgroups.
constexpr int kMsHeavyRow = 256;
// Builds the degree-ordered copy of `g` (msbfs.hip); called by the
// DeviceGraph constructor with the device current.
void build_ms_graph(DeviceGraph& dg, const HostGraph& g);
bool ms_config_supported(int block, int k);
// HBM a batch workspace needs beyond the graph (plus the staging buffer when
// results go to the host).
size_t batch_bytes_estimate(int64_t n, int64_t m, int k, bool host_levels);

// Coordinate-descent tuners (autotune.cpp): leave the fastest configuration
// on the handle.
void autotune_bfs(DeviceGraph& g, const int32_t* sources, int count, int reps, gb_tune_result* out);
void autotune_batch(DeviceGraph& g, const int32_t* sources, int k, int reps, gb_tune_result* out);

// Device-wide exclusive scan of int32 (scan.hip).
void exclusive_scan_i32(int device, const int32_t* d_in, int32_t* d_out, int64_t n);
// The same scan with its partials in a fixed device array: no allocation,
// at most exclusive_scan_noalloc_max() items; overlapping calls on one GPU
// share that array and race.
void exclusive_scan_i32_noalloc(int device, const int32_t* d_in, int32_t* d_out, int64_t n);
int64_t exclusive_scan_noalloc_max();

}  // namespace gg
