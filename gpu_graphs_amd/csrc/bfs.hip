// Single-source top-down BFS for gfx950.
//
// Replaces the reference's level loop (1D Partition/src/gpuBFS.cu:66-100) and
// its kernels `linear_bfs` (gpuBFS_kernels.cu:16-80, thread per queue entry,
// serial neighbour loop, one global atomicAdd per edge, no visited filter) and
// the unlaunched degree-tiered `expand_contract_bfs` (gpuBFS_kernels.cu:
// 249-478).  Level semantics follow the reference CPU oracle cpuBFS
// (1D Partition/src/cpuBFS.cpp:34-58): a vertex first seen while expanding
// level L gets level L+1.
//
// Design (one GPU, whole graph resident in HBM):
//  * Frontier entries are (first CSR entry, exclusive degree prefix, vertex).
//    Entries get their queue slot and their edge offset from ONE 64-bit
//    atomicAdd per workgroup tile on a packed (count << 32 | edges) counter,
//    so the degree prefix needed for edge-parallel work distribution comes
//    out of the enqueue itself -- no separate scan pass per level.
//  * A level is cut into tiles of T = BLOCK*K consecutive frontier edges.
//    The enqueuer records, for every tile boundary inside a new entry's edge
//    range, which entry owns it (tile_start), so a workgroup finds its
//    entries with one load.  Every thread then owns K edges of the tile
//    (coalesced column reads), whatever the degrees: hubs with 10^5 edges are
//    spread over many workgroups instead of one thread.
//  * Visited claims: plain load of the bitmap word (may be stale, never
//    wrongly "visited"), then atomicOr; the winner writes level L+1 (and the
//    parent).  Levels are therefore deterministic.
//  * Winners are appended to an LDS list with wave-aggregated LDS atomics
//    (ballot + mbcnt), scanned once per tile (block scan) and written out.
//  * Persistent mode runs all levels in one cooperative launch with a grid
//    barrier between levels; level-loop mode launches once per level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "device_graph.hpp"
#include "wave.hpp"

namespace gg {

namespace {

// ctl[] layout (64-bit words).  Every word that workgroups use to agree with
// each other (ring counters, barrier, abort) is read and written only with
// atomic read-modify-write operations, which are coherent across the XCDs'
// separate L2 caches; plain loads of these words could return a stale line.
enum : int {
    CTL_RING0 = 0,      // 3 packed frontier counters: level L uses ring[L % 3]
    CTL_BAR_COUNT = 3,  // grid barrier arrivals, monotonic (never reset in a launch)
    CTL_ABORT = 4,      // set when a barrier wait times out; every workgroup then exits
    CTL_NREACH = 5,     // reached vertices
    CTL_MREACH = 6,     // sum of degrees of reached vertices
    CTL_MAXLEV = 7,     // largest level assigned
    CTL_ERR = 8,        // internal consistency error (non-zero = bad)
    CTL_BAR_SEEN = 9,   // on timeout: arrivals observed by the workgroup that gave up
    CTL_BAR_WANT = 10,  // on timeout: arrivals it waited for
    CTL_WORDS = 16
};

enum : unsigned long long { ERR_TILE = 1, ERR_LEVELS = 2 };

struct BfsArgs {
    const int64_t* __restrict__ offsets;
    const int32_t* __restrict__ cols;
    uint32_t* visited;
    int32_t* levels;
    int32_t* parents;
    int64_t* q_rb[2];
    uint32_t* q_eoff[2];
    int32_t* q_v[2];
    uint32_t* tile_start[2];
    unsigned long long* ctl;
    int64_t n;
    uint32_t nwords;
    int32_t source;
    uint64_t spin_limit;  // wall-clock ticks a barrier may wait before aborting
};

template <int BLOCK, int K>
struct TileLds {
    static constexpr int T = BLOCK * K;
    uint32_t eoff[T + 2];     // edge prefix of the tile's entries (+ end sentinel)
    int64_t rb[T + 1];        // first CSR entry of each tile entry
    int32_t v[T + 1];         // vertex id of each tile entry (parents only)
    int64_t win_rb[T];        // winners with degree > 0: first CSR entry
    uint32_t win_deg[T + 1];  // their degree, then (in place) exclusive prefix
    int32_t win_v[T];         // their vertex id
    uint32_t wave_tmp[BLOCK / 64];
    uint32_t nwin;
    uint32_t nreach;
    unsigned long long base;
};

// Coherent read of a control word (atomic add of zero).
__device__ __forceinline__ unsigned long long ctl_read(unsigned long long* p) {
    return __hip_atomic_fetch_add(p, 0ull, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier over all workgroups of a cooperative launch.  The arrival
// counter only grows: barrier number `epoch` (0, 1, ...) is complete when it
// reaches (epoch + 1) * nblocks, so no workgroup reads anything before it
// arrives and a stale read can only delay, never release, a waiter.  The
// wait is bounded by `spin_limit` wall-clock ticks: on timeout the waiter
// raises CTL_ABORT (recording what it saw) and every workgroup leaves the
// BFS loop, so a launch that is not fully co-resident ends with an error
// instead of hanging.  Returns false when the BFS must stop.
__device__ bool grid_barrier(const BfsArgs& a, uint32_t nblocks, unsigned long long epoch) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release this workgroup's queue / level / tile_start stores
        const unsigned long long target = (epoch + 1) * nblocks;
        unsigned long long seen = __hip_atomic_fetch_add(&a.ctl[CTL_BAR_COUNT], 1ull, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_AGENT) + 1;
        int ok = 1;
        const uint64_t t0 = wall_clock64();
        while (seen < target) {
            __builtin_amdgcn_s_sleep(8);
            seen = ctl_read(&a.ctl[CTL_BAR_COUNT]);
            if (seen >= target) break;
            if (ctl_read(&a.ctl[CTL_ABORT])) {
                ok = 0;
                break;
            }
            if (static_cast<uint64_t>(wall_clock64()) - t0 > a.spin_limit) {
                if (atomicOr(&a.ctl[CTL_ABORT], 1ull) == 0) {
                    atomicExch(&a.ctl[CTL_BAR_SEEN], seen);
                    atomicExch(&a.ctl[CTL_BAR_WANT], target);
                }
                ok = 0;
                break;
            }
        }
        __threadfence();  // acquire the other workgroups' stores
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Expands tile t of level L: frontier edges [t*T, min((t+1)*T, E)).
template <int BLOCK, int K>
__device__ void expand_tile(const BfsArgs& a, TileLds<BLOCK, K>& s, int L, uint32_t t, uint32_t ntiles,
                            uint32_t F, uint32_t E) {
    constexpr uint32_t T = BLOCK * K;
    const int cur = L & 1, nxt = cur ^ 1;
    const uint32_t tid = threadIdx.x;
    const uint32_t base = t * T;
    const uint64_t tile_end = uint64_t(base) + T;
    const uint32_t end = tile_end < E ? static_cast<uint32_t>(tile_end) : E;
    const uint32_t i0 = a.tile_start[cur][t];
    const uint32_t i1 = (t + 1 < ntiles) ? a.tile_start[cur][t + 1] : F - 1;
    uint32_t c = i1 - i0 + 1;
    if (i1 < i0 || c > T + 1 || i1 >= F) {  // cannot happen; guard the LDS arrays
        if (tid == 0) atomicOr(&a.ctl[CTL_ERR], ERR_TILE);
        c = 0;
    }
    for (uint32_t j = tid; j <= c; j += BLOCK) {
        if (j < c) {
            s.eoff[j] = a.q_eoff[cur][i0 + j];
            s.rb[j] = a.q_rb[cur][i0 + j];
            if (a.parents) s.v[j] = a.q_v[cur][i0 + j];
        } else {
            s.eoff[j] = (i0 + c < F) ? a.q_eoff[cur][i0 + c] : E;
        }
    }
    if (tid == 0) {
        s.nwin = 0;
        s.nreach = 0;
    }
    __syncthreads();
    if (c == 0) return;

    // Gather: K independent column loads per thread, then K bitmap probes.
    int32_t u[K];
    uint32_t owner[K];
    bool act[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t e = base + k * BLOCK + tid;
        act[k] = e < end;
        u[k] = 0;
        owner[k] = 0;
        if (act[k]) {
            uint32_t lo = 0, hi = c - 1;  // last entry with eoff <= e
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (s.eoff[mid] <= e) lo = mid;
                else hi = mid - 1;
            }
            owner[k] = lo;
            u[k] = a.cols[s.rb[lo] + (e - s.eoff[lo])];
        }
    }
    uint32_t word[K];
#pragma unroll
    for (int k = 0; k < K; ++k) word[k] = act[k] ? a.visited[static_cast<uint32_t>(u[k]) >> 5] : ~0u;

    const uint32_t lane = tid & 63u;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t bit = 1u << (static_cast<uint32_t>(u[k]) & 31u);
        bool win = false;
        int64_t rb = 0;
        uint32_t deg = 0;
        if (act[k] && !(word[k] & bit)) {
            uint32_t old = atomicOr(&a.visited[static_cast<uint32_t>(u[k]) >> 5], bit);
            if (!(old & bit)) {
                win = true;
                a.levels[u[k]] = L + 1;
                if (a.parents) a.parents[u[k]] = s.v[owner[k]];
                rb = a.offsets[u[k]];
                deg = static_cast<uint32_t>(a.offsets[u[k] + 1] - rb);
            }
        }
        const uint64_t any = wave_ballot(win);
        if (any) {
            const uint64_t keep = wave_ballot(win && deg > 0);
            if (lane == 0) atomicAdd(&s.nreach, static_cast<uint32_t>(__popcll(any)));
            if (keep) {
                uint32_t pos = 0;
                if (lane == 0) pos = atomicAdd(&s.nwin, static_cast<uint32_t>(__popcll(keep)));
                pos = __shfl(pos, 0, 64);
                if (win && deg > 0) {
                    const uint32_t slot = pos + rank_below(keep);
                    s.win_rb[slot] = rb;
                    s.win_deg[slot] = deg;
                    s.win_v[slot] = u[k];
                }
            }
        }
    }
    __syncthreads();

    const uint32_t nwin = s.nwin;
    if (tid == 0 && s.nreach) {
        atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(s.nreach));
        atomicMax(&a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L + 1));
    }
    if (nwin == 0) {
        __syncthreads();
        return;
    }
    // Exclusive prefix of winner degrees, in place, BLOCK entries per round.
    uint32_t running = 0;
    for (uint32_t r0 = 0; r0 < nwin; r0 += BLOCK) {
        const uint32_t i = r0 + tid;
        const uint32_t d = (i < nwin) ? s.win_deg[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<BLOCK>(d, s.wave_tmp, tot);
        if (i < nwin) s.win_deg[i] = running + ex;
        running += tot;
    }
    if (tid == 0) {
        s.win_deg[nwin] = running;
        s.base = atomicAdd(&a.ctl[CTL_RING0 + (L + 1) % 3],
                           (static_cast<unsigned long long>(nwin) << 32) | running);
    }
    __syncthreads();
    const uint32_t slot0 = static_cast<uint32_t>(s.base >> 32);
    const uint32_t e0 = static_cast<uint32_t>(s.base);
    for (uint32_t i = tid; i < nwin; i += BLOCK) {
        const uint32_t slot = slot0 + i;
        const uint32_t eo = e0 + s.win_deg[i];
        const uint32_t d = s.win_deg[i + 1] - s.win_deg[i];
        a.q_rb[nxt][slot] = s.win_rb[i];
        a.q_eoff[nxt][slot] = eo;
        if (a.parents) a.q_v[nxt][slot] = s.win_v[i];
        const uint64_t stop = uint64_t(eo) + d;
        for (uint64_t tt = (uint64_t(eo) + T - 1) / T; tt * T < stop; ++tt) a.tile_start[nxt][tt] = slot;
    }
    __syncthreads();
}

// Reads the packed frontier counter of level L once per workgroup and
// broadcasts it through LDS, so all threads of a workgroup agree.  After an
// abort it reads as empty, so a late-starting workgroup does nothing.
__device__ __forceinline__ unsigned long long level_counter(const BfsArgs& a, int L) {
    __shared__ unsigned long long s_cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        s_cnt = ctl_read(&a.ctl[CTL_RING0 + L % 3]);
        if (ctl_read(&a.ctl[CTL_ABORT])) s_cnt = 0;
    }
    __syncthreads();
    return s_cnt;
}

template <int BLOCK, int K>
__device__ void expand_level(const BfsArgs& a, TileLds<BLOCK, K>& s, int L, unsigned long long cnt, uint32_t b,
                             uint32_t G) {
    constexpr uint32_t T = BLOCK * K;
    const uint32_t F = static_cast<uint32_t>(cnt >> 32), E = static_cast<uint32_t>(cnt);
    if (b == 0 && threadIdx.x == 0) {
        // ring[(L+2)%3] held level L-1's count, read by every workgroup
        // before the barrier that ended level L-1; it next counts level L+2.
        atomicExch(&a.ctl[CTL_RING0 + (L + 2) % 3], 0ull);
        atomicAdd(&a.ctl[CTL_MREACH], static_cast<unsigned long long>(E));
    }
    if (F == 0) return;
    const uint32_t ntiles = static_cast<uint32_t>((uint64_t(E) + T - 1) / T);
    for (uint32_t t = b; t < ntiles; t += G) expand_tile<BLOCK, K>(a, s, L, t, ntiles, F, E);
}

template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void bfs_level_kernel(BfsArgs a, int L) {
    __shared__ TileLds<BLOCK, K> s;
    expand_level<BLOCK, K>(a, s, L, level_counter(a, L), blockIdx.x, gridDim.x);
}

// All levels in one cooperative launch.  A BFS has at most n levels; the
// loop is also bounded by that, so a corrupted counter cannot spin forever.
template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void bfs_persistent_kernel(BfsArgs a) {
    __shared__ TileLds<BLOCK, K> s;
    for (int L = 0;; ++L) {
        const unsigned long long cnt = level_counter(a, L);
        if ((cnt >> 32) == 0) break;  // every workgroup reads the same final value
        if (L > a.n) {
            if (threadIdx.x == 0) atomicOr(&a.ctl[CTL_ERR], ERR_LEVELS);
            break;
        }
        expand_level<BLOCK, K>(a, s, L, cnt, blockIdx.x, gridDim.x);
        if (!grid_barrier(a, gridDim.x, static_cast<unsigned long long>(L))) break;
    }
}

// levels = -1, parents = -1, visited = 0.
__global__ void bfs_clear_kernel(BfsArgs a) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        a.levels[i] = -1;
        if (a.parents) a.parents[i] = -1;
        if (i < a.nwords) a.visited[i] = 0u;
    }
}

// Level 0: the source alone (claimed, level 0, parent = itself).
__global__ void bfs_seed_kernel(BfsArgs a, uint32_t T) {
    const int32_t s = a.source;
    const int64_t rb = a.offsets[s];
    const uint32_t d = static_cast<uint32_t>(a.offsets[s + 1] - rb);
    if (threadIdx.x == 0) {
        a.visited[static_cast<uint32_t>(s) >> 5] = 1u << (static_cast<uint32_t>(s) & 31u);
        a.levels[s] = 0;
        if (a.parents) a.parents[s] = s;
        a.q_rb[0][0] = rb;
        a.q_eoff[0][0] = 0;
        a.q_v[0][0] = s;
        for (int i = 0; i < CTL_WORDS; ++i) a.ctl[i] = 0;
        a.ctl[CTL_RING0] = d ? ((1ull << 32) | d) : 0ull;
        a.ctl[CTL_NREACH] = 1;
    }
    for (uint64_t tt = threadIdx.x; tt * T < d; tt += blockDim.x) a.tile_start[0][tt] = 0;
}

using PersistentFn = void (*)(BfsArgs);
using LevelFn = void (*)(BfsArgs, int);

struct KernelSet {
    int block, k;
    PersistentFn persistent;
    LevelFn level;
};

#define GG_KSET(B, K) {B, K, bfs_persistent_kernel<B, K>, bfs_level_kernel<B, K>}
const KernelSet kKernels[] = {GG_KSET(256, 1), GG_KSET(256, 2), GG_KSET(256, 4), GG_KSET(256, 8),
                              GG_KSET(512, 1), GG_KSET(512, 2), GG_KSET(512, 4), GG_KSET(1024, 1),
                              GG_KSET(1024, 2), GG_KSET(128, 4), GG_KSET(128, 8), GG_KSET(64, 8),
                              GG_KSET(64, 16)};
#undef GG_KSET

const KernelSet* find_kernels(int block, int k) {
    for (const auto& ks : kKernels)
        if (ks.block == block && ks.k == k) return &ks;
    return nullptr;
}

}  // namespace

bool bfs_config_supported(int block, int k) { return find_kernels(block, k) != nullptr; }

DeviceGraph::DeviceGraph(const HostGraph& g, int dev) : device(dev), n(g.n), m(g.m()) {
    if (m >= (int64_t(1) << 32) - 1) fail(GB_EINVAL, "graphs with 2^32 or more CSR entries are not supported");
    if (n < 1) fail(GB_EINVAL, "graph has no vertices");
    int count = 0;
    GG_HIP(hipGetDeviceCount(&count));
    if (dev < 0 || dev >= count) fail(GB_EINVAL, "device " + std::to_string(dev) + " does not exist");
    DeviceGuard guard(device);
    hipDeviceProp_t prop;
    GG_HIP(hipGetDeviceProperties(&prop, device));
    num_cus = prop.multiProcessorCount;
    int clock_khz = 0;
    GG_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, device));
    if (clock_khz <= 0) clock_khz = 100000;
    spin_limit_ticks = static_cast<uint64_t>(clock_khz) * 1000ull * kBarrierTimeoutSeconds;
    try {
        GG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        GG_HIP(hipEventCreate(&ev_start));
        GG_HIP(hipEventCreate(&ev_stop));
        GG_HIP(hipEventCreate(&ev_kernel));
        GG_HIP(hipMalloc(&offsets, sizeof(int64_t) * (n + 1)));
        GG_HIP(hipMalloc(&cols, sizeof(int32_t) * std::max<int64_t>(m, 1)));
        graph_bytes = sizeof(int64_t) * (n + 1) + sizeof(int32_t) * std::max<int64_t>(m, 1);
        GG_HIP(hipMemcpyAsync(offsets, g.offsets.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, stream));
        if (m) GG_HIP(hipMemcpyAsync(cols, g.cols.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, stream));
        GG_HIP(hipStreamSynchronize(stream));
    } catch (...) {
        this->~DeviceGraph();
        throw;
    }
}

DeviceGraph::~DeviceGraph() {
    (void)hipSetDevice(device);
    free_workspace();
    if (offsets) (void)hipFree(offsets);
    if (cols) (void)hipFree(cols);
    offsets = nullptr;
    cols = nullptr;
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (ev_kernel) (void)hipEventDestroy(ev_kernel);
    if (stream) (void)hipStreamDestroy(stream);
    ev_start = ev_stop = ev_kernel = nullptr;
    stream = nullptr;
}

void DeviceGraph::free_workspace() {
    void* ptrs[] = {ws.visited, ws.q_rb[0], ws.q_rb[1], ws.q_eoff[0], ws.q_eoff[1], ws.q_v[0], ws.q_v[1],
                    ws.tile_start[0], ws.tile_start[1], ws.ctl, ws.levels, ws.parents};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    ws = BfsWorkspace{};
}

void DeviceGraph::ensure_workspace() {
    const int T = cfg.block * cfg.edges_per_thread;
    const int64_t tile_cap = (m + T - 1) / T + 2;
    if (ws.ctl && ws.tile_edges == T && ws.tile_cap >= tile_cap) return;
    if (ws.ctl && ws.tile_edges != T) {  // only the tile index depends on T
        if (ws.tile_start[0]) (void)hipFree(ws.tile_start[0]);
        if (ws.tile_start[1]) (void)hipFree(ws.tile_start[1]);
        ws.bytes -= 2 * sizeof(uint32_t) * ws.tile_cap;
        ws.tile_start[0] = ws.tile_start[1] = nullptr;
        GG_HIP(hipMalloc(&ws.tile_start[0], sizeof(uint32_t) * tile_cap));
        GG_HIP(hipMalloc(&ws.tile_start[1], sizeof(uint32_t) * tile_cap));
        ws.tile_cap = tile_cap;
        ws.tile_edges = T;
        ws.bytes += 2 * sizeof(uint32_t) * tile_cap;
        return;
    }
    const int64_t nwords = (n + 31) / 32;
    size_t bytes = 0;
    auto alloc = [&](auto** p, size_t sz) {
        GG_HIP(hipMalloc(reinterpret_cast<void**>(p), sz));
        bytes += sz;
    };
    try {
        alloc(&ws.visited, sizeof(uint32_t) * nwords);
        for (int i = 0; i < 2; ++i) {
            alloc(&ws.q_rb[i], sizeof(int64_t) * n);
            alloc(&ws.q_eoff[i], sizeof(uint32_t) * n);
            alloc(&ws.q_v[i], sizeof(int32_t) * n);
            alloc(&ws.tile_start[i], sizeof(uint32_t) * tile_cap);
        }
        alloc(&ws.ctl, sizeof(unsigned long long) * CTL_WORDS);
        alloc(&ws.levels, sizeof(int32_t) * n);
        alloc(&ws.parents, sizeof(int32_t) * n);
    } catch (...) {
        free_workspace();
        throw;
    }
    ws.tile_cap = tile_cap;
    ws.tile_edges = T;
    ws.bytes = bytes;
}

void DeviceGraph::bfs(int32_t source, int32_t* d_levels, int32_t* d_parents, gb_stats* st) {
    if (source < 0 || source >= n)
        fail(GB_ERANGE, "source " + std::to_string(source) + " outside [0, " + std::to_string(n) + ")");
    if (!d_levels) fail(GB_EINVAL, "levels buffer is null");
    const KernelSet* ks = find_kernels(cfg.block, cfg.edges_per_thread);
    if (!ks) fail(GB_EINVAL, "unsupported block / edges_per_thread combination");
    DeviceGuard guard(device);
    ensure_workspace();
    const uint32_t T = static_cast<uint32_t>(cfg.block * cfg.edges_per_thread);

    BfsArgs a;
    a.offsets = offsets;
    a.cols = cols;
    a.visited = ws.visited;
    a.levels = d_levels;
    a.parents = d_parents;
    for (int i = 0; i < 2; ++i) {
        a.q_rb[i] = ws.q_rb[i];
        a.q_eoff[i] = ws.q_eoff[i];
        a.q_v[i] = ws.q_v[i];
        a.tile_start[i] = ws.tile_start[i];
    }
    a.ctl = ws.ctl;
    a.n = n;
    a.nwords = static_cast<uint32_t>((n + 31) / 32);
    a.source = source;
    a.spin_limit = spin_limit_ticks;

    int occ = 0;
    if (cfg.mode == GB_MODE_PERSISTENT)
        GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(ks->persistent),
                                                           cfg.block, 0));
    else
        GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(ks->level),
                                                           cfg.block, 0));
    if (occ < 1) fail(GB_EDEVICE, "BFS kernel cannot be resident on a CU");
    // Default: at most 16 waves per CU (half the wave slots), leaving
    // headroom so every workgroup of the cooperative launch is co-resident.
    const int waves_per_block = (cfg.block + 63) / 64;
    const int default_per_cu = std::max(1, std::min(occ, 16 / waves_per_block));
    const int per_cu = cfg.blocks_per_cu > 0 ? std::min(cfg.blocks_per_cu, occ) : default_per_cu;
    const uint32_t grid = static_cast<uint32_t>(per_cu * num_cus);
    const uint32_t clear_grid = static_cast<uint32_t>(std::min<int64_t>((n + 255) / 256, int64_t(num_cus) * 8));

    GG_HIP(hipEventRecord(ev_start, stream));
    hipLaunchKernelGGL(bfs_clear_kernel, dim3(clear_grid), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(bfs_seed_kernel, dim3(1), dim3(256), 0, stream, a, T);
    GG_HIP(hipGetLastError());
    GG_HIP(hipEventRecord(ev_kernel, stream));
    if (cfg.mode == GB_MODE_PERSISTENT) {
        void* args[] = {&a};
        GG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(ks->persistent), dim3(grid), dim3(cfg.block),
                                          args, 0, stream));
    } else {
        unsigned long long* h_cnt = nullptr;
        GG_HIP(hipHostMalloc(&h_cnt, sizeof(unsigned long long) * CTL_WORDS, hipHostMallocDefault));
        try {
            unsigned long long cnt = 1ull << 32;
            for (int64_t L = 0; (cnt >> 32) != 0; ++L) {
                if (L > n) fail(GB_EDEVICE, "internal error: more BFS levels than vertices");
                const uint32_t E = L == 0 ? 0u : static_cast<uint32_t>(cnt);
                uint32_t g = grid;
                if (L > 0) g = static_cast<uint32_t>(std::min<uint64_t>(grid, (uint64_t(E) + T - 1) / T));
                hipLaunchKernelGGL(ks->level, dim3(std::max(1u, g)), dim3(cfg.block), 0, stream, a, static_cast<int>(L));
                GG_HIP(hipGetLastError());
                GG_HIP(hipMemcpyAsync(h_cnt, ws.ctl, sizeof(unsigned long long) * CTL_WORDS, hipMemcpyDeviceToHost,
                                      stream));
                GG_HIP(hipStreamSynchronize(stream));
                cnt = h_cnt[CTL_RING0 + (L + 1) % 3];
            }
        } catch (...) {
            (void)hipHostFree(h_cnt);
            throw;
        }
        GG_HIP(hipHostFree(h_cnt));
    }
    GG_HIP(hipEventRecord(ev_stop, stream));
    unsigned long long h_ctl[CTL_WORDS];
    GG_HIP(hipMemcpyAsync(h_ctl, ws.ctl, sizeof(h_ctl), hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
    if (h_ctl[CTL_ABORT])
        fail(GB_EDEVICE, "persistent BFS aborted: a grid barrier waited longer than the limit (" +
                             std::to_string(h_ctl[CTL_BAR_SEEN]) + " of " + std::to_string(h_ctl[CTL_BAR_WANT]) +
                             " arrivals seen; " + std::to_string(grid) + " workgroups of " +
                             std::to_string(cfg.block) + " threads were not all co-resident)");
    if (h_ctl[CTL_ERR] & ERR_TILE) fail(GB_EDEVICE, "internal error: inconsistent frontier tile index");
    if (h_ctl[CTL_ERR] & ERR_LEVELS) fail(GB_EDEVICE, "internal error: more BFS levels than vertices");
    if (st) {
        float ms = 0.f, kms = 0.f;
        GG_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
        GG_HIP(hipEventElapsedTime(&kms, ev_kernel, ev_stop));
        st->ms = ms;
        st->kernel_ms = kms;
        st->num_levels = static_cast<int32_t>(h_ctl[CTL_MAXLEV]) + 1;
        st->reserved = 0;
        st->n_reached = static_cast<int64_t>(h_ctl[CTL_NREACH]);
        st->m_reached = static_cast<int64_t>(h_ctl[CTL_MREACH]);
    }
}

void DeviceGraph::bfs_to_host(int32_t source, int32_t* h_levels, int32_t* h_parents, gb_stats* st) {
    if (!h_levels) fail(GB_EINVAL, "levels buffer is null");
    DeviceGuard guard(device);
    ensure_workspace();
    bfs(source, ws.levels, h_parents ? ws.parents : nullptr, st);
    GG_HIP(hipMemcpyAsync(h_levels, ws.levels, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
    if (h_parents) GG_HIP(hipMemcpyAsync(h_parents, ws.parents, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
}

}  // namespace gg
