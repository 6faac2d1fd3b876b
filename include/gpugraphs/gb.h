/*
 * gb.h -- C ABI of the MI355X (gfx950) BFS engine.
 *
 * This is the drop-in boundary for the single-source BFS path of
 * YKMorsy/GPU_Graphs.  The reference has no FFI layer: its surface is three
 * C++ classes built by `main` (1D Partition/main.cu:17-35):
 *
 *   csr::csr(const char* filename)      1D Partition/include/csr.h:12
 *       public int* row_offset, int* col_idx, int num_nodes, int num_edges
 *                                        1D Partition/include/csr.h:17-20
 *   gpuBFS::gpuBFS(csr&, int source)     1D Partition/include/gpuBFS.cuh:13
 *       public int* host_distance, float exec_time, int iteration,
 *       int total_edges_traversed        1D Partition/include/gpuBFS.cuh:17-20
 *   PrefixSum::sum_scan_blelloch(int* d_out, const int* d_in, size_t n)
 *                                        1D Partition/include/prefixSum.cuh:7-9
 *
 * Each entry point below names the reference interface it replaces.  The C++
 * facade in gpugraphs.hpp rebuilds the three classes (same member names) on
 * top of these functions, and the Python package gpu_graphs_amd binds them
 * with ctypes.
 *
 * Conventions
 *   - Every function returns GB_OK (0) or a negative GB_E* status.  The
 *     library never calls exit() and no C++ exception crosses the ABI.  The
 *     message for the last failing call on the calling thread is returned by
 *     gb_last_error().
 *   - Vertex ids are 0-based int32.  Levels are int32 with -1 for "not
 *     reached" (the reference uses INT_MAX, 1D Partition/include/cpuBFS.h:11).
 *   - Output buffers are caller-owned.  Functions with a _device suffix take
 *     device pointers (memory on the graph's GPU) and leave results in HBM;
 *     the others take host pointers and include the device-to-host copy.
 *   - Handles are library-owned and released with the matching *_free.
 *     A gb_dgraph is not thread-safe; distinct handles may be used from
 *     distinct threads.
 */
#ifndef GPUGRAPHS_GB_H
#define GPUGRAPHS_GB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GB_OK        0
#define GB_EINVAL   -1   /* bad argument (null handle, k out of range, ...) */
#define GB_EIO      -2   /* file missing, unreadable or malformed          */
#define GB_ENOMEM   -3   /* host or device allocation failed               */
#define GB_EDEVICE  -4   /* HIP runtime error                              */
#define GB_ERANGE   -5   /* vertex id / source outside [0, n)              */
#define GB_EBUSY    -6   /* server: memory budget or queue full            */
#define GB_EAGAIN   -7   /* server: result not ready yet                   */

/* ---- graph construction flags ----------------------------------------- */
#define GB_ONE_BASED        1u  /* ids in the file are 1-based (MIG loader, csr.cpp:28 `from-1,to-1`) */
#define GB_SYMMETRIZE       2u  /* add the reverse of every edge                      */
#define GB_KEEP_ORDER       4u  /* rows keep file order, duplicates and self-loops,
                                   exactly as csr::csr lays them out (csr.cpp:32-57) */
#define GB_DROP_SELF_LOOPS  8u  /* remove u->u edges                                  */
#define GB_AUTO_N          16u  /* n = max(header n, largest id + 1): for files whose ids
                                   exceed the header (graphs/wiki-Vote.txt, which the
                                   reference loader overflows)                       */

typedef struct gb_graph  gb_graph;   /* host CSR: int64 offsets, int32 columns */
typedef struct gb_dgraph gb_dgraph;  /* device-resident copy on one GPU        */
typedef struct gb_server gb_server;  /* continuous-batching BFS query server   */

/* Per-BFS statistics.  n_reached / m_reached follow the Graph500 counting
 * used for GTEPS: m_reached is the number of directed CSR entries of reached
 * vertices (GTEPS uses m_reached/2).  The reference's own counters are broken
 * (always 0: 1D Partition/src/gpuBFS_kernels.cu:37, gpuBFS.cu:63); these are
 * exact. */
typedef struct gb_stats {
    double  ms;           /* device time, level init + source injection -> last
                           * level: GPU clock stamps in the first and last kernel
                           * (GB_MODE_ASYNC), hipEvents (the other modes)         */
    double  kernel_ms;    /* traversal kernel(s) alone, without the level init
                           * (GB_MODE_LEVEL_LOOP; the other modes report ms)      */
    int32_t num_levels;   /* 1 + max level                                           */
    int32_t reserved;
    int64_t n_reached;
    int64_t m_reached;
    double  copy_ms;      /* gb_bfs: device-to-host copy of levels (+ parents), the
                           * reference's get_device_distance (gpuBFS.cu:118-137),
                           * hipEvent time outside `ms`; 0 for gb_bfs_device       */
} gb_stats;

/* ---- errors ------------------------------------------------------------ */
const char* gb_last_error(void);
const char* gb_version(void);

/* ---- host graph (replaces csr::csr, 1D Partition/src/csr.cpp:3-67) ----- */

/* Reads the reference edge-list format: line 1 `n`, line 2 `m` (a hint; the
 * reference trusts it, csr.cpp:16-22), then whitespace-separated `from to`
 * pairs.  Ids outside [0,n) after the GB_ONE_BASED shift are GB_ERANGE (the
 * reference overflows, csr.cpp:28). */
int  gb_graph_load_edgelist(const char* path, uint32_t flags, gb_graph** out);
/* Builds a CSR from an edge list held in memory (src[i] -> dst[i]). */
int  gb_graph_from_edges(int64_t n, int64_t m, const int32_t* src, const int32_t* dst,
                         uint32_t flags, gb_graph** out);
/* Copies an existing CSR (offsets[n+1], cols[offsets[n]]); rows are sorted
 * and deduplicated unless GB_KEEP_ORDER. */
int  gb_graph_from_csr(int64_t n, const int64_t* offsets, const int32_t* cols,
                       uint32_t flags, gb_graph** out);
/* Synthetic inputs.  All are symmetric, without self-loops or duplicates,
 * and deterministic for a given seed. */
int  gb_graph_generate_rmat(int scale, int edge_factor, double a, double b, double c,
                            uint64_t seed, int permute_ids, gb_graph** out);
int  gb_graph_generate_er(int64_t n_candidates, int64_t n_isolated, int64_t m_undirected,
                          uint64_t seed, gb_graph** out);
int  gb_graph_generate_grid(int64_t width, int64_t height, double extra_fraction,
                            uint64_t seed, gb_graph** out);
/* n = num_nodes, m = num_edges (directed CSR entries). */
int  gb_graph_info(const gb_graph* g, int64_t* n, int64_t* m);
/* Borrowed views of row_offset / col_idx (valid until gb_graph_free). */
int  gb_graph_csr(const gb_graph* g, const int64_t** offsets, const int32_t** cols);
/* Degree-sorted vertex buckets, computed on the first call: bucket 0 holds the
 * vertices of degree 0, bucket b >= 1 those with degree in [2^(b-1), 2^b).
 * Fills offsets[GB_DEGREE_BUCKETS + 1]; bucket b is (*vertices)[offsets[b] ..
 * offsets[b+1]), ids ascending inside a bucket; *vertices is borrowed (n
 * entries, valid until gb_graph_free).  Replaces the degree tiers that the
 * reference computes per frontier entry (gpuBFS_kernels.cu:451-478). */
#define GB_DEGREE_BUCKETS 34
int  gb_graph_degree_buckets(gb_graph* g, int64_t* offsets, const int32_t** vertices);
/* *symmetric = 1 when every edge has its reverse (built with GB_SYMMETRIZE, by
 * a generator, or found so on sorted rows); such graphs may be searched
 * bottom-up. */
int  gb_graph_is_symmetric(const gb_graph* g, int* symmetric);
/* Writes the graph in the reference text format (one `u v` line per CSR entry). */
int  gb_graph_write_edgelist(const gb_graph* g, const char* path);
/* Binary CSR cache (little-endian header + offsets + cols). */
int  gb_graph_save_binary(const gb_graph* g, const char* path);
int  gb_graph_load_binary(const char* path, gb_graph** out);
void gb_graph_free(gb_graph* g);

/* ---- device graph (replaces gpuBFS::init_device, gpuBFS.cu:139-185) ---- */
int  gb_device_count(int* count);
int  gb_device_upload(const gb_graph* g, int device, gb_dgraph** out);
int  gb_dgraph_info(const gb_dgraph* dg, int64_t* n, int64_t* m, int* device);
/* Bytes of HBM the handle holds (graph + single-source workspace). */
int  gb_dgraph_memory(const gb_dgraph* dg, int64_t* bytes);
void gb_device_free(gb_dgraph* dg);

/* ---- single-source BFS (replaces gpuBFS::gpuBFS, gpuBFS.cu:13-116) ----- */

/* Execution modes for gb_bfs_config. */
#define GB_MODE_ASYNC       0   /* levels queued ahead in batches, no per-level host sync
                                   (default)                                               */
#define GB_MODE_LEVEL_LOOP  1   /* host reads the frontier size after every level        */
#define GB_MODE_GRAPH       2   /* as ASYNC, the first levels replayed from one captured
                                   hipGraph                                                */

/* levels: n int32 (host), -1 = unreached.  parents: nullable, n int32,
 * parent[source] = source, -1 = unreached (any valid BFS tree). */
int  gb_bfs(gb_dgraph* dg, int32_t source, int32_t* levels, int32_t* parents, gb_stats* st);
/* Same, with levels / parents in device memory of the graph's GPU. */
int  gb_bfs_device(gb_dgraph* dg, int32_t source, int32_t* d_levels, int32_t* d_parents,
                   gb_stats* st);
/* Tunables: expand() workgroup size and edges per thread per tile (supported
 * pairs: 256x{1,2,4,8}, 512x{1,2,4}, 1024x{1,2}, 128x{4,8}, 64x{8,16}),
 * workgroups per CU (0 = at most 16 waves per CU), mode, and the frontier
 * edge count above which a level's next frontier is rebuilt from the
 * visited bitmap by a scan instead of queued by atomics (0 = default,
 * max(2^19, m/256)). */
int  gb_bfs_config(gb_dgraph* dg, int block, int edges_per_thread, int blocks_per_cu,
                   int mode, int64_t dense_edges);
int  gb_bfs_get_config(const gb_dgraph* dg, int* block, int* edges_per_thread,
                       int* blocks_per_cu, int* mode, int64_t* dense_edges);
/* Direction optimisation of gb_bfs (bottom-up levels, Beamer, Asanovic and
 * Patterson, SC 2012), used on symmetric graphs only: enable (1/0); alpha:
 * go bottom-up when the frontier grows and its edges * alpha > unexplored
 * edges (default 14); beta: back to top-down when the next frontier shrinks
 * and has fewer than n / beta vertices (default 24).
 * alpha or beta 0 keeps the current value.  The reference has no
 * counterpart (top-down only, gpuBFS.cu:66-100). */
int  gb_bfs_direction_config(gb_dgraph* dg, int enable, int alpha, int beta);
int  gb_bfs_get_direction_config(const gb_dgraph* dg, int* enable, int* alpha, int* beta, int* symmetric);
/* Deep mode of gb_bfs (no reference counterpart; the reference pays a host
 * round trip per level, gpuBFS.cu:66-100): one launch per level of a kernel
 * built for thin frontiers -- a thread per frontier vertex, claims straight
 * into the visited bitmap, one queue reservation per wave, top-down only --
 * and the level's counters are kept by that kernel.  GB_DEEP_AUTO
 * (default) uses it once searches of the graph have been deep (>= 64 levels)
 * with thin frontiers only, such as road networks; GB_DEEP_ON always;
 * GB_DEEP_OFF never.  *used (nullable) tells whether the last gb_bfs ran in
 * deep mode. */
#define GB_DEEP_AUTO 0
#define GB_DEEP_ON   1
#define GB_DEEP_OFF  2
int  gb_bfs_deep_config(gb_dgraph* dg, int mode);
int  gb_bfs_get_deep_config(const gb_dgraph* dg, int* mode, int* used);
/* The tail of a search (no reference counterpart): on symmetric graphs
 * outside deep mode, the launches stop at a level learned per handle (the
 * latest start of a final run of small levels in recent searches), from
 * which one workgroup runs the rest of the search with workgroup barriers.
 * A level too large for it goes back to the per-level kernels.  level:
 * GB_TAIL_AUTO (default, learned), GB_TAIL_OFF, or a level >= 0 to start
 * the tail there.  Results do not depend on it.  *used (nullable): level the
 * last gb_bfs started its tail at, -1 if none; *outcome (nullable):
 * GB_TAIL_FINISHED, GB_TAIL_DECLINED (its first level was too large; the
 * per-level kernels went on from there), GB_TAIL_HANDED_BACK (a later level
 * was too large; the deep-mode kernel went on from there), 0 if no tail. */
#define GB_TAIL_AUTO (-1)
#define GB_TAIL_OFF  (-2)
#define GB_TAIL_FINISHED    1
#define GB_TAIL_DECLINED    2
#define GB_TAIL_HANDED_BACK 3
int  gb_bfs_tail_config(gb_dgraph* dg, int level);
int  gb_bfs_get_tail_config(const gb_dgraph* dg, int* level, int* used, int* outcome);
/* Levels the last gb_bfs on this handle searched bottom-up (-1, -1 if none). */
int  gb_bfs_last_bottom_up(const gb_dgraph* dg, int* first_level, int* last_level);
/* Launch hints learned from the searches so far: levels launched before the
 * first done check, and the levels [bu_lo, bu_hi] that get a bottom-up
 * launch (bu_hi < bu_lo: none). */
int  gb_bfs_launch_hints(const gb_dgraph* dg, int* depth, int* bu_lo, int* bu_hi);

/* ---- batched multi-source BFS (no reference counterpart: the reference
 *      runs one source per gpuBFS object, 1D Partition/main.cu:17-35).
 *      Up to 64 sources share one launch sequence with 64-bit masks; each
 *      source's levels equal gb_bfs's for that source. ------------------- */
typedef struct gb_batch_stats {
    double  ms;              /* device time of the whole batch (hipEvent)       */
    int32_t num_levels;      /* 1 + max level over the batch                    */
    int32_t k;
    int64_t n_reached[64];   /* per source                                       */
    int64_t m_reached[64];   /* per source: sum of degrees of reached vertices   */
} gb_batch_stats;

/* sources: k host ints (1 <= k <= 64, repeats allowed).  levels: k rows of n
 * int32, row i = levels from sources[i].  st (nullable) adds a statistics
 * pass after the timed region. */
int  gb_bfs_batch(gb_dgraph* dg, const int32_t* sources, int k, int32_t* levels, gb_batch_stats* st);
int  gb_bfs_batch_device(gb_dgraph* dg, const int32_t* sources, int k, int32_t* d_levels,
                         gb_batch_stats* st);
/* Batch kernel shape (same pairs as gb_bfs_config), whether a batch is
 * replayed as one captured hipGraph (1) or launched kernel by kernel (0), and
 * the frontier size (CSR entries) up to which a level hands a list of the
 * vertices it reached to the next one instead of a scan of all n vertices
 * (0 = default, max(2^16, n/16)). */
int  gb_batch_config(gb_dgraph* dg, int block, int edges_per_thread, int use_graph, int64_t sparse_edges);
int  gb_batch_get_config(const gb_dgraph* dg, int* block, int* edges_per_thread, int* use_graph,
                         int64_t* sparse_edges);
/* Bottom-up levels of gb_bfs_batch (symmetric graphs only): enable (1/0);
 * alpha: a level goes bottom-up when its frontier edges * alpha exceed the
 * edges of the vertices not yet reached by every source (default 2; 0 keeps
 * the current value); beta: and only while the frontier holds at least
 * (vertices with edges) / beta vertices (0: no such limit, the default; -1
 * keeps the current value). */
int  gb_batch_direction_config(gb_dgraph* dg, int enable, int alpha, int beta);
int  gb_batch_get_direction_config(const gb_dgraph* dg, int* enable, int* alpha, int* beta);
/* Number of levels the last gb_bfs_batch on this handle ran bottom-up. */
int  gb_batch_last_bottom_up(const gb_dgraph* dg, int* levels);
/* Deep mode of gb_bfs_batch: one fused launch per level (each vertex of the
 * level's list takes its new bits and pushes them along its row), top-down
 * only.  GB_DEEP_AUTO (default) uses it once a batch on the graph has been
 * deep (>= 64 levels) with thin frontiers only; GB_DEEP_ON always;
 * GB_DEEP_OFF never.  *used (nullable): whether the last batch used it. */
int  gb_batch_deep_config(gb_dgraph* dg, int mode);
int  gb_batch_get_deep_config(const gb_dgraph* dg, int* mode, int* used);

/* ---- autotuner (no reference counterpart: the reference's kernel shape is
 *      the compile-time BLOCK_SIZE 1024, 1D Partition/include/
 *      gpuBFS_kernels.cuh, and it has no thresholds).  Coordinate descent
 *      over the kernel shape, the direction thresholds, the bottom-up
 *      degree tier and the frontier-size threshold, each candidate timed on
 *      the device (best of `reps` passes after a warm-up); the fastest
 *      configuration is left on the handle.  Results are exact for every
 *      candidate; only the time changes. ------------------------------- */
typedef struct gb_tune_result {
    double  default_ms;        /* configuration on entry: ms per call       */
    double  best_ms;           /* chosen configuration: ms per call         */
    int32_t tried;             /* configurations timed                      */
    int32_t block;             /* chosen workgroup size                     */
    int32_t edges_per_thread;  /* chosen edges per thread per tile          */
    int32_t alpha;             /* chosen top-down -> bottom-up threshold    */
    int32_t beta;              /* chosen bottom-up -> top-down threshold (single-source only, else 0) */
    int32_t low_degree;        /* batch bottom-up: rows up to this long go to one lane (0 for gb_bfs) */
    int64_t frontier_edges;    /* gb_bfs: dense_edges; gb_bfs_batch: sparse_edges */
} gb_tune_result;

/* Tunes gb_bfs over `count` (1..64) sources: each pass runs them all. */
int  gb_bfs_autotune(gb_dgraph* dg, const int32_t* sources, int count, int reps, gb_tune_result* out);
/* Tunes gb_bfs_batch on one batch of k (1..64) sources. */
int  gb_batch_autotune(gb_dgraph* dg, const int32_t* sources, int k, int reps, gb_tune_result* out);
/* Applies a tuning result (from this or an earlier process) to a handle.
 * default_ms, best_ms and tried are ignored; beta 0 keeps the current one;
 * gb_bfs_apply_tuning ignores low_degree. */
int  gb_bfs_apply_tuning(gb_dgraph* dg, const gb_tune_result* t);
int  gb_batch_apply_tuning(gb_dgraph* dg, const gb_tune_result* t);

/* ---- continuous-batching query server (no reference counterpart: the
 *      reference answers one (graph, source) per process run, 1D Partition/
 *      main.cu:10-82).  Queries on any added graph are batched (up to
 *      max_batch per graph) onto `lanes` concurrent GPU streams as lanes and
 *      device memory free up; admission is by a device-memory budget that
 *      covers lane workspaces and result buffers.  Results stay in device
 *      memory until gb_server_release.  The gb_dgraph handles must outlive
 *      the server.  All gb_server_* calls are thread-safe. -------------- */
typedef struct gb_server_cfg {
    int     device;
    int     max_batch;       /* 1..64 (0 = 64)                                    */
    int     lanes;           /* concurrent batches, 1..16 (0 = 4)                 */
    int     plain;           /* 1: one single-source BFS per query, no batching   */
    int64_t memory_budget;   /* bytes (0 = half of the free device memory)        */
    int64_t queue_capacity;  /* queued queries before submit returns GB_EBUSY (0 = 2^20) */
} gb_server_cfg;

typedef struct gb_latency {
    double queue_ms;   /* arrival -> batch start */
    double run_ms;     /* batch start -> done    */
    double total_ms;   /* arrival -> done        */
    int    batch_size;
} gb_latency;

typedef struct gb_server_stats {
    int64_t submitted, completed, failed, rejected, batches;
    double  mean_batch;
    double  p50_ms, p99_ms, max_ms;   /* arrival -> done of completed queries */
    double  qps;                      /* completed / (last done - first arrival) */
    int64_t memory_used, memory_budget;
    int64_t workspace_drops;          /* lane workspaces freed to admit other graphs' batches */
} gb_server_stats;

int  gb_server_create(const gb_server_cfg* cfg, gb_server** out);
int  gb_server_add_graph(gb_server* srv, gb_dgraph* dg, int* gid);
/* GB_EBUSY when the queue is full, GB_ERANGE for a bad source. */
int  gb_server_submit(gb_server* srv, int gid, int32_t source, uint64_t* ticket);
/* GB_OK when done (levels, if not null, receives n int32 on the host),
 * GB_EAGAIN while pending, or the query's failure status (GB_ENOMEM when its
 * batch could never fit the budget, GB_EDEVICE on a device error). */
int  gb_server_poll(gb_server* srv, uint64_t ticket, int32_t* levels, gb_latency* lat);
int  gb_server_wait(gb_server* srv, uint64_t ticket, int timeout_ms, int32_t* levels, gb_latency* lat);
/* Device pointer of a finished query's n levels (valid until release). */
int  gb_server_result_device(gb_server* srv, uint64_t ticket, const int32_t** d_levels);
/* Frees a finished query's result (or cancels a queued one). */
int  gb_server_release(gb_server* srv, uint64_t ticket);
/* Waits until nothing is queued or running (GB_EAGAIN on timeout). */
int  gb_server_drain(gb_server* srv, int timeout_ms);
int  gb_server_get_stats(gb_server* srv, gb_server_stats* st);
int  gb_server_reset_stats(gb_server* srv);
void gb_server_free(gb_server* srv);

/* ---- device-wide exclusive scan (replaces PrefixSum::sum_scan_blelloch,
 *      1D Partition/src/prefixSum.cu:139-224).  d_out[i] = sum(d_in[0..i)).
 *      In-place (d_out == d_in) is allowed.  No allocation per call. ----- */
int  gb_exclusive_scan_i32(int device, const int32_t* d_in, int32_t* d_out, int64_t n);

/* ---- allocation-free exclusive scan.  No counterpart in the reference: it
 *      is for callers that cannot allocate device memory (code under HIP
 *      graph capture, tight serving loops, hosts that forbid hipMalloc).
 *      Same result as gb_exclusive_scan_i32, but it calls no hipMalloc,
 *      hipFree or allocating library routine: its per-tile partials live in
 *      a fixed 1 MiB device array, one per GPU, held until the process ends.
 *      n may be at most gb_exclusive_scan_noalloc_max() (2^30); longer
 *      inputs fail with GB_EINVAL.  Calls that overlap on one GPU share the
 *      array and race: serialise them per device. ----------------------- */
int  gb_exclusive_scan_i32_noalloc(int device, const int32_t* d_in, int32_t* d_out, int64_t n);
int64_t gb_exclusive_scan_noalloc_max(void);

/* ---- device memory helpers (so hosts need no other GPU runtime) -------- */
int  gb_dmalloc(int device, size_t bytes, void** ptr);
int  gb_dfree(int device, void* ptr);
int  gb_memcpy_htod(int device, void* dst, const void* src, size_t bytes);
int  gb_memcpy_dtoh(int device, void* dst, const void* src, size_t bytes);
int  gb_device_synchronize(int device);

#ifdef __cplusplus
}
#endif
#endif /* GPUGRAPHS_GB_H */
