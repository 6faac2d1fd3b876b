// Single-source top-down BFS for gfx950.
//
// Replaces the reference's level loop (1D Partition/src/gpuBFS.cu:66-100) and
// its kernels `linear_bfs` (gpuBFS_kernels.cu:16-80: thread per queue entry,
// serial neighbour loop, one global atomicAdd per edge, no visited filter) and
// the unlaunched degree-tiered `expand_contract_bfs` (gpuBFS_kernels.cu:
// 249-478).  Level semantics follow the reference CPU oracle cpuBFS
// (1D Partition/src/cpuBFS.cpp:34-58): a vertex first seen while expanding
// level L gets level L+1.
//
// Design (one GPU, whole graph resident in HBM).  Measured on MI355X: one
// same-address 64-bit atomic costs ~12 ns (they serialise), a random
// atomicOr into an L2-resident bitmap ~0.04 ns of chip time, back-to-back
// kernel launches ~1.5 us, a software grid barrier 17-57 us.  Hence no grid
// barrier and no per-tile atomics on a shared counter in large levels:
//
//  * A level L runs as two launches, `expand(L)` then `compact(L)`, issued
//    back to back with no host synchronisation; the host only checks a done
//    flag once per batch of levels.
//  * The frontier of a level is a queue of entries (vertex, first CSR entry,
//    exclusive prefix of degrees).  expand(L) cuts the level's edges into
//    tiles of T = BLOCK*K edges; tile_start[t] names the entry owning edge
//    t*T, so each workgroup loads its entries with one lookup and every
//    thread owns K edges of the tile (coalesced column reads) whatever the
//    degree distribution: hubs are spread over many workgroups.
//  * Claims: plain load of the visited-bitmap word (may be stale, never
//    wrongly "visited"), then atomicOr; the winner writes level L+1 and the
//    parent.  Levels are deterministic.
//  * Sparse levels (at most `dense_edges` frontier edges) build the next
//    queue directly: winners go to an LDS list (wave-aggregated LDS atomics,
//    ballot + mbcnt), get a block scan, and ONE 64-bit atomicAdd per tile on
//    a packed (count << 32 | edges) counter gives both their queue slots and
//    their edge offsets.
//  * Dense levels skip the enqueue.  compact(L) rebuilds the queue from the
//    bitmaps: new = visited & ~prev over 4096-vertex chunks -- per-thread
//    serial sums, a wave64 scan and an LDS-staged block scan give each new
//    vertex its offset inside the chunk, and the chunk reserves its queue
//    slots and edge range with one packed 64-bit atomicAdd, like a sparse
//    tile.  No workgroup waits on another, so the kernels need no particular
//    residency and can share the device with other streams' work.
//  * Every kernel stops at once when the error word is set, and the expand
//    tiles check each queue entry and column index against n and m, so an
//    inconsistent queue is reported as an error, never dereferenced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "device_graph.hpp"
#include "frontier.hpp"
#include "wave.hpp"

namespace gg {

namespace {

// ctl[] layout (64-bit words).
enum : int {
    CTL_RING0 = 0,   // 3 packed (count << 32 | edges) frontier counters: level L uses ring[L % 3]
    CTL_NREACH = 3,  // reached vertices
    CTL_MREACH = 4,  // sum of degrees of reached vertices
    CTL_MAXLEV = 5,  // largest level assigned
    CTL_DONE = 6,    // set by the first expand() that finds an empty frontier
    CTL_ERR = 7,     // internal consistency error (bit set)
    CTL_WORDS = 16
};
enum : unsigned long long { ERR_TILE = 1, ERR_RANGE = 2 };

constexpr int kCompactBlock = 256;
constexpr int kCompactRounds = 16;                          // vertices per thread
constexpr int kChunk = kCompactBlock * kCompactRounds;      // vertices per compaction chunk

struct BfsArgs {
    const int64_t* __restrict__ offsets;
    const int32_t* __restrict__ cols;
    uint32_t* vis;    // claimed vertices
    uint32_t* prev;   // vertices already placed in (or handled by) a frontier queue
    int32_t* levels;
    int32_t* parents;
    int64_t* q_rb[2];
    uint32_t* q_eoff[2];
    int32_t* q_v[2];
    uint32_t* tile_start[2];
    unsigned long long* ctl;
    int64_t n;
    int64_t m;
    uint32_t nwords;
    uint32_t nchunks;
    uint32_t dense_edges;
    uint32_t tile_edges;        // T of the expand kernel (tile_start granularity)
    int32_t source;
};

template <int BLOCK, int K>
struct TileLds {
    static constexpr int T = BLOCK * K;
    uint32_t eoff[T + 2];     // edge prefix of the tile's entries (+ end sentinel)
    int64_t rb[T + 1];        // first CSR entry of each tile entry
    int32_t v[T + 1];         // vertex id of each tile entry
    int64_t win_rb[T];        // sparse levels: winners with degree > 0
    uint32_t win_deg[T + 1];  // their degree, then (in place) exclusive prefix
    int32_t win_v[T];
    uint32_t wave_tmp[BLOCK / 64];
    uint32_t nwin;
    uint32_t nzero;
    unsigned long long base;
};

// Expands tile t of level L: frontier edges [t*T, min((t+1)*T, E)).
template <int BLOCK, int K>
__device__ void expand_tile(const BfsArgs& a, TileLds<BLOCK, K>& s, int L, uint32_t t, uint32_t ntiles,
                            uint32_t F, uint32_t E, bool dense) {
    constexpr uint32_t T = BLOCK * K;
    const int cur = L & 1, nxt = cur ^ 1;
    const uint32_t tid = threadIdx.x;
    const uint32_t base = t * T;
    const uint64_t tile_end = uint64_t(base) + T;
    const uint32_t end = tile_end < E ? static_cast<uint32_t>(tile_end) : E;
    const uint32_t i0 = a.tile_start[cur][t];
    const uint32_t i1 = (t + 1 < ntiles) ? a.tile_start[cur][t + 1] : F - 1;
    uint32_t c = i1 - i0 + 1;
    if (i1 < i0 || c > T + 1 || i1 >= F) {  // cannot happen; guards the LDS arrays
        if (tid == 0) atomicOr(&a.ctl[CTL_ERR], ERR_TILE);
        c = 0;
    }
    for (uint32_t j = tid; j <= c; j += BLOCK) {
        if (j < c) {
            s.eoff[j] = a.q_eoff[cur][i0 + j];
            s.rb[j] = a.q_rb[cur][i0 + j];
            s.v[j] = a.q_v[cur][i0 + j];
        } else {
            s.eoff[j] = (i0 + c < F) ? a.q_eoff[cur][i0 + c] : E;
        }
    }
    if (tid == 0) {
        s.nwin = 0;
        s.nzero = 0;
    }
    __syncthreads();
    if (c == 0) return;

    // Gather: K independent column loads per thread, then K bitmap probes.
    // Edge e belongs to entry lo iff eoff[lo] <= e < eoff[lo + 1]; entries and
    // columns that fail the range checks are dropped and reported.
    int32_t u[K];
    uint32_t owner[K];
    bool act[K];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t e = base + k * BLOCK + tid;
        act[k] = e < end;
        u[k] = 0;
        owner[k] = 0;
        if (act[k]) {
            const uint32_t lo = owner_of(s.eoff, c, e);
            owner[k] = lo;
            const uint64_t pos = static_cast<uint64_t>(s.rb[lo]) + (e - s.eoff[lo]);
            if (e < s.eoff[lo] || e >= s.eoff[lo + 1] || pos >= static_cast<uint64_t>(a.m)) {
                act[k] = false;
                bad = true;
            } else {
                u[k] = a.cols[pos];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (act[k] && static_cast<uint32_t>(u[k]) >= static_cast<uint64_t>(a.n)) {
            act[k] = false;
            bad = true;
        }
    }
    if (bad) atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
    uint32_t word[K];
#pragma unroll
    for (int k = 0; k < K; ++k) word[k] = act[k] ? a.vis[static_cast<uint32_t>(u[k]) >> 5] : ~0u;

    // Claims: all K atomics are issued before any result is used, so their
    // round trips overlap instead of running back to back.
    uint32_t old[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t uu = static_cast<uint32_t>(u[k]);
        const uint32_t bit = 1u << (uu & 31u);
        old[k] = (act[k] && !(word[k] & bit)) ? atomicOr(&a.vis[uu >> 5], bit) : ~0u;
    }
    bool won[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t uu = static_cast<uint32_t>(u[k]);
        won[k] = !(old[k] & (1u << (uu & 31u)));
        if (won[k]) {
            a.levels[uu] = L + 1;
            if (a.parents) a.parents[uu] = s.v[owner[k]];
        }
    }
    if (dense) {  // compact() builds the next queue from the bitmaps
        __syncthreads();
        return;
    }
    int64_t wrb[K];
    int64_t wre[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t uu = static_cast<uint32_t>(u[k]);
        wrb[k] = wre[k] = 0;
        if (won[k]) {
            atomicOr(&a.prev[uu >> 5], 1u << (uu & 31u));  // queued (or degree 0): compact() must skip it
            wrb[k] = a.offsets[uu];
            wre[k] = a.offsets[uu + 1];
        }
    }
    const uint32_t lane = tid & 63u;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t rb = wrb[k];
        const uint32_t deg = static_cast<uint32_t>(wre[k] - wrb[k]);
        const bool win = won[k];
        const uint64_t keep = wave_ballot(win && deg > 0);
        const uint64_t zero = wave_ballot(win && deg == 0);
        if (zero && lane == 0) atomicAdd(&s.nzero, static_cast<uint32_t>(__popcll(zero)));
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
    __syncthreads();

    const uint32_t nwin = s.nwin;
    if (tid == 0 && s.nzero) {  // degree-0 winners are reached but never queued
        atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(s.nzero));
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
        const unsigned long long tot = (static_cast<unsigned long long>(nwin) << 32) | running;
        s.base = atomicAdd(&a.ctl[CTL_RING0 + (L + 1) % 3], tot);
        if (!reservation_fits(s.base, tot, a.n, a.m)) {
            atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
            s.nwin = 0;  // nothing is written
        }
    }
    __syncthreads();
    if (s.nwin == 0) {
        __syncthreads();
        return;
    }
    const uint32_t slot0 = static_cast<uint32_t>(s.base >> 32);
    const uint32_t e0 = static_cast<uint32_t>(s.base);
    for (uint32_t i = tid; i < nwin; i += BLOCK) {
        const uint32_t slot = slot0 + i;
        const uint32_t eo = e0 + s.win_deg[i];
        const uint32_t d = s.win_deg[i + 1] - s.win_deg[i];
        a.q_rb[nxt][slot] = s.win_rb[i];
        a.q_eoff[nxt][slot] = eo;
        a.q_v[nxt][slot] = s.win_v[i];
        mark_tiles(a.tile_start[nxt], T, eo, d, slot);
    }
    __syncthreads();
}

// Level L, step 1: expand the frontier queue of level L.
template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void bfs_expand_kernel(BfsArgs a, int L) {
    __shared__ TileLds<BLOCK, K> s;
    constexpr uint32_t T = BLOCK * K;
    // Levels queued past the end of the search are no-ops: once done, the
    // ring slots are no longer reset and would hold stale counts.  After an
    // error the queue is not trusted: nothing runs.
    if (a.ctl[CTL_DONE] | a.ctl[CTL_ERR]) return;
    const unsigned long long cnt = a.ctl[CTL_RING0 + L % 3];  // final: written by earlier launches
    const uint32_t F = static_cast<uint32_t>(cnt >> 32), E = static_cast<uint32_t>(cnt);
    if (F == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[CTL_DONE] = 1;
        return;
    }
    if (F > a.n || E > a.m) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // ring[(L+2)%3] held level L-1's counter; it next counts level L+2.
        a.ctl[CTL_RING0 + (L + 2) % 3] = 0;
        atomicAdd(&a.ctl[CTL_MREACH], static_cast<unsigned long long>(E));
        atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(F));
        atomicMax(&a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L));
    }
    const bool dense = E > a.dense_edges;
    const uint32_t ntiles = static_cast<uint32_t>((uint64_t(E) + T - 1) / T);
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) expand_tile<BLOCK, K>(a, s, L, t, ntiles, F, E, dense);
}

// Level L, step 2 (dense levels only): queue of level L+1 from the bitmaps.
// Chunk c covers vertices [c*kChunk, (c+1)*kChunk); thread t takes vertices
// c*kChunk + r*256 + t for r < 16, so the offsets loads of a wave are
// contiguous.  The thread's new vertices occupy consecutive queue slots; the
// chunk's block goes where its atomicAdd on level L+1's counter puts it.
__device__ void compact_chunk(const BfsArgs& a, int L, uint32_t c) {
    __shared__ unsigned long long s_wave[kCompactBlock / 64];
    __shared__ unsigned long long s_excl;
    __shared__ uint32_t s_ok;
    __shared__ uint32_t s_zero[kCompactBlock / 64];
    const uint32_t tid = threadIdx.x;
    const int64_t vbase = int64_t(c) * kChunk;
    uint32_t newmask = 0;  // bit r: vertex vbase + r*256 + tid is new
    uint32_t kept = 0, zero = 0, edges = 0;
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        const int64_t v = vbase + r * kCompactBlock + tid;
        if (v < a.n) {
            const uint32_t w = static_cast<uint32_t>(v >> 5);
            const uint32_t nb = a.vis[w] & ~a.prev[w];
            if ((nb >> (v & 31)) & 1u) {
                const uint32_t d = static_cast<uint32_t>(a.offsets[v + 1] - a.offsets[v]);
                if (d) {
                    newmask |= 1u << r;
                    ++kept;
                    edges += d;
                } else {
                    ++zero;
                }
            }
        }
    }
    // Block scan of (kept, edges) packed in 64 bits; zero-degree count summed.
    const unsigned long long mine = (static_cast<unsigned long long>(kept) << 32) | edges;
    unsigned long long total;
    const unsigned long long excl_local = block_exclusive_scan<kCompactBlock>(mine, s_wave, total);
    const uint32_t zsum = wave_sum(zero);
    if ((tid & 63) == 0) s_zero[tid >> 6] = zsum;
    // prev |= vis for the chunk's words: every read of prev above happened
    // before the block scan's barriers.
    for (uint32_t i = tid; i < kChunk / 32; i += kCompactBlock) {
        const int64_t w = vbase / 32 + i;
        if (w < a.nwords) a.prev[w] = a.vis[w];
    }
    if (tid == 0) {
        s_ok = 1;
        if (total) {
            s_excl = atomicAdd(&a.ctl[CTL_RING0 + (L + 1) % 3], total);
            if (!reservation_fits(s_excl, total, a.n, a.m)) {
                atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
                s_ok = 0;
            }
        }
        uint32_t z = 0;
        for (int i = 0; i < kCompactBlock / 64; ++i) z += s_zero[i];
        if (z) {
            atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(z));
            atomicMax(&a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L + 1));
        }
    }
    __syncthreads();
    if (!newmask || !s_ok) return;
    const unsigned long long pre = s_excl + excl_local;
    uint32_t slot = static_cast<uint32_t>(pre >> 32);
    uint32_t eo = static_cast<uint32_t>(pre);
    const int q = (L + 1) & 1;
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        if ((newmask >> r) & 1u) {
            const int64_t v = vbase + r * kCompactBlock + tid;
            const int64_t rb = a.offsets[v];
            const uint32_t d = static_cast<uint32_t>(a.offsets[v + 1] - rb);
            a.q_rb[q][slot] = rb;
            a.q_eoff[q][slot] = eo;
            a.q_v[q][slot] = static_cast<int32_t>(v);
            mark_tiles(a.tile_start[q], a.tile_edges, eo, d, slot);
            ++slot;
            eo += d;
        }
    }
}

// Level L+1's counter ring[(L+1)%3] was zeroed by expand(L-1) and, on a
// dense level, expand(L) left it alone: the chunks' reservations build it up.
__global__ __launch_bounds__(kCompactBlock) void bfs_compact_kernel(BfsArgs a, int L) {
    if (a.ctl[CTL_DONE] | a.ctl[CTL_ERR]) return;
    const unsigned long long cnt = a.ctl[CTL_RING0 + L % 3];
    const uint32_t F = static_cast<uint32_t>(cnt >> 32), E = static_cast<uint32_t>(cnt);
    if (F == 0 || E <= a.dense_edges) return;  // sparse level: expand() queued level L+1
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
        compact_chunk(a, L, c);
        __syncthreads();
    }
}

// levels = -1, parents = -1, bitmaps = 0.
__global__ void bfs_clear_kernel(BfsArgs a) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        a.levels[i] = -1;
        if (a.parents) a.parents[i] = -1;
        if (i < a.nwords) {
            a.vis[i] = 0u;
            a.prev[i] = 0u;
        }
    }
}

// Level 0: the source alone (claimed, level 0, parent = itself).
__global__ void bfs_seed_kernel(BfsArgs a) {
    const int32_t s = a.source;
    const int64_t rb = a.offsets[s];
    const uint32_t d = static_cast<uint32_t>(a.offsets[s + 1] - rb);
    if (threadIdx.x == 0) {
        const uint32_t bit = 1u << (static_cast<uint32_t>(s) & 31u);
        a.vis[static_cast<uint32_t>(s) >> 5] = bit;
        a.prev[static_cast<uint32_t>(s) >> 5] = bit;
        a.levels[s] = 0;
        if (a.parents) a.parents[s] = s;
        a.q_rb[0][0] = rb;
        a.q_eoff[0][0] = 0;
        a.q_v[0][0] = s;
        for (int i = 0; i < CTL_WORDS; ++i) a.ctl[i] = 0;
        a.ctl[CTL_RING0] = d ? ((1ull << 32) | d) : 0ull;
        a.ctl[CTL_NREACH] = d ? 0 : 1;  // expand(0) counts a source with edges
    }
    for (uint64_t tt = threadIdx.x; tt * a.tile_edges < d; tt += blockDim.x) a.tile_start[0][tt] = 0;
}

using ExpandFn = void (*)(BfsArgs, int);

struct KernelSet {
    int block, k;
    ExpandFn expand;
};

#define GG_KSET(B, K) {B, K, bfs_expand_kernel<B, K>}
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
    if (n >= (int64_t(1) << 30)) fail(GB_EINVAL, "graphs with 2^30 or more vertices are not supported");
    int count = 0;
    GG_HIP(hipGetDeviceCount(&count));
    if (dev < 0 || dev >= count) fail(GB_EINVAL, "device " + std::to_string(dev) + " does not exist");
    DeviceGuard guard(device);
    hipDeviceProp_t prop;
    GG_HIP(hipGetDeviceProperties(&prop, device));
    num_cus = prop.multiProcessorCount;
    cfg.dense_edges = default_dense_edges();
    try {
        GG_HIP(hipMalloc(&offsets, sizeof(int64_t) * (n + 1)));
        GG_HIP(hipMalloc(&cols, sizeof(int32_t) * std::max<int64_t>(m, 1)));
        graph_bytes = sizeof(int64_t) * (n + 1) + sizeof(int32_t) * std::max<int64_t>(m, 1);
        GG_HIP(hipMemcpy(offsets, g.offsets.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
        if (m) GG_HIP(hipMemcpy(cols, g.cols.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
        runner = std::make_unique<Runner>(*this);
    } catch (...) {
        this->~DeviceGraph();
        throw;
    }
}

DeviceGraph::~DeviceGraph() {
    (void)hipSetDevice(device);
    runner.reset();
    if (offsets) (void)hipFree(offsets);
    if (cols) (void)hipFree(cols);
    offsets = nullptr;
    cols = nullptr;
}

Runner::Runner(DeviceGraph& graph) : g(graph), device(graph.device), n(graph.n), m(graph.m) {
    DeviceGuard guard(device);
    try {
        GG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        GG_HIP(hipEventCreate(&ev_start));
        GG_HIP(hipEventCreate(&ev_stop));
        GG_HIP(hipEventCreate(&ev_kernel));
        GG_HIP(hipEventCreateWithFlags(&ev_batch[0], hipEventDisableTiming));
        GG_HIP(hipEventCreateWithFlags(&ev_batch[1], hipEventDisableTiming));
        GG_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_ctl), sizeof(unsigned long long) * 2 * kCtlWords,
                             hipHostMallocDefault));
    } catch (...) {
        this->~Runner();
        throw;
    }
}

Runner::~Runner() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    free_workspace();
    free_ms_workspace();
    if (h_ctl) (void)hipHostFree(h_ctl);
    h_ctl = nullptr;
    for (hipEvent_t* e : {&ev_start, &ev_stop, &ev_kernel, &ev_batch[0], &ev_batch[1]}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

void Runner::free_workspace() {
    void* ptrs[] = {ws.vis,    ws.prev,    ws.q_rb[0],  ws.q_rb[1],       ws.q_eoff[0],    ws.q_eoff[1],
                    ws.q_v[0], ws.q_v[1],  ws.ctl,      ws.levels,        ws.parents,      ws.tile_start[0],
                    ws.tile_start[1]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    ws = BfsWorkspace{};
}

void Runner::ensure_workspace() {
    const int T = g.cfg.block * g.cfg.edges_per_thread;
    const int64_t tile_cap = (m + T - 1) / T + 2;
    if (ws.ctl && ws.tile_edges == T) return;
    if (ws.ctl) {  // only the tile index depends on T
        (void)hipFree(ws.tile_start[0]);
        (void)hipFree(ws.tile_start[1]);
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
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    size_t bytes = 0;
    auto alloc = [&](auto** p, size_t sz) {
        GG_HIP(hipMalloc(reinterpret_cast<void**>(p), sz));
        bytes += sz;
    };
    try {
        alloc(&ws.vis, sizeof(uint32_t) * nwords);
        alloc(&ws.prev, sizeof(uint32_t) * nwords);
        for (int i = 0; i < 2; ++i) {
            alloc(&ws.q_rb[i], sizeof(int64_t) * n);
            alloc(&ws.q_eoff[i], sizeof(uint32_t) * n);
            alloc(&ws.q_v[i], sizeof(int32_t) * n);
            alloc(&ws.tile_start[i], sizeof(uint32_t) * tile_cap);
        }
        alloc(&ws.ctl, sizeof(unsigned long long) * kCtlWords);
        alloc(&ws.levels, sizeof(int32_t) * n);
        alloc(&ws.parents, sizeof(int32_t) * n);
    } catch (...) {
        free_workspace();
        throw;
    }
    ws.tile_cap = tile_cap;
    ws.tile_edges = T;
    ws.nchunks = nchunks;
    ws.bytes = bytes;
}

void Runner::bfs(int32_t source, int32_t* d_levels, int32_t* d_parents, gb_stats* st) {
    static_assert(kCtlWords == CTL_WORDS, "control block size");
    if (source < 0 || source >= n)
        fail(GB_ERANGE, "source " + std::to_string(source) + " outside [0, " + std::to_string(n) + ")");
    if (!d_levels) fail(GB_EINVAL, "levels buffer is null");
    const KernelSet* ks = find_kernels(g.cfg.block, g.cfg.edges_per_thread);
    if (!ks) fail(GB_EINVAL, "unsupported block / edges_per_thread combination");
    DeviceGuard guard(device);
    ensure_workspace();

    BfsArgs a;
    a.offsets = g.offsets;
    a.cols = g.cols;
    a.vis = ws.vis;
    a.prev = ws.prev;
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
    a.m = m;
    a.nwords = static_cast<uint32_t>((n + 31) / 32);
    a.nchunks = static_cast<uint32_t>(ws.nchunks);
    a.dense_edges = static_cast<uint32_t>(std::min<int64_t>(g.cfg.dense_edges, 0xffffffffll));
    a.tile_edges = static_cast<uint32_t>(g.cfg.block * g.cfg.edges_per_thread);
    a.source = source;

    int occ = 0;
    GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(ks->expand), g.cfg.block, 0));
    if (occ < 1) fail(GB_EDEVICE, "BFS kernel cannot be resident on a CU");
    const int waves_per_block = (g.cfg.block + 63) / 64;
    const int default_per_cu = std::max(1, std::min(occ, 16 / waves_per_block));
    const int per_cu = g.cfg.blocks_per_cu > 0 ? std::min(g.cfg.blocks_per_cu, occ) : default_per_cu;
    const uint32_t grid = static_cast<uint32_t>(per_cu * g.num_cus);
    // compact(): chunks are independent; 4 workgroups of 256 per CU.
    const uint32_t cgrid = static_cast<uint32_t>(std::min<int64_t>(ws.nchunks, int64_t(g.num_cus) * 4));
    const uint32_t clear_grid = static_cast<uint32_t>(std::min<int64_t>((n + 255) / 256, int64_t(g.num_cus) * 8));

    auto launch_level = [&](int L) {
        hipLaunchKernelGGL(ks->expand, dim3(grid), dim3(g.cfg.block), 0, stream, a, L);
        hipLaunchKernelGGL(bfs_compact_kernel, dim3(cgrid), dim3(kCompactBlock), 0, stream, a, L);
    };

    GG_HIP(hipEventRecord(ev_start, stream));
    hipLaunchKernelGGL(bfs_clear_kernel, dim3(clear_grid), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(bfs_seed_kernel, dim3(1), dim3(256), 0, stream, a);
    GG_HIP(hipGetLastError());
    GG_HIP(hipEventRecord(ev_kernel, stream));
    int L = 0;
    if (g.cfg.mode == GB_MODE_LEVEL_LOOP) {
        // One level at a time; the host reads the next frontier size.
        for (;; ++L) {
            if (L > n) fail(GB_EDEVICE, "internal error: more BFS levels than vertices");
            launch_level(L);
            GG_HIP(hipGetLastError());
            GG_HIP(hipMemcpyAsync(h_ctl, ws.ctl, sizeof(unsigned long long) * kCtlWords, hipMemcpyDeviceToHost,
                                  stream));
            GG_HIP(hipStreamSynchronize(stream));
            if (h_ctl[CTL_ERR]) break;
            if ((h_ctl[CTL_RING0 + (L + 1) % 3] >> 32) == 0) {
                ++L;
                launch_level(L);  // sets the done flag and the final counters
                break;
            }
        }
    } else {
        // Levels are queued ahead with no host synchronisation.  The first
        // batch covers the previous BFS's depth + 1 (one empty level ends the
        // search); if the search is deeper, batches of doubling size (at most
        // 64 levels) are pipelined: the done flag of batch b-1 is checked
        // while batch b runs.  Extra levels are no-ops on an empty frontier.
        auto enqueue = [&](int count, int slot) {
            for (int i = 0; i < count; ++i) launch_level(L++);
            GG_HIP(hipGetLastError());
            GG_HIP(hipMemcpyAsync(h_ctl + slot * kCtlWords, ws.ctl, sizeof(unsigned long long) * kCtlWords,
                                  hipMemcpyDeviceToHost, stream));
            GG_HIP(hipEventRecord(ev_batch[slot], stream));
        };
        auto done = [&](int slot) {
            GG_HIP(hipEventSynchronize(ev_batch[slot]));
            return h_ctl[slot * kCtlWords + CTL_DONE] != 0 || h_ctl[slot * kCtlWords + CTL_ERR] != 0;
        };
        enqueue(last_num_levels > 0 ? last_num_levels + 1 : 8, 0);
        if (!done(0)) {
            int batch = 8;
            enqueue(batch, 1);
            for (int b = 2;; ++b) {
                if (L > n + 64) fail(GB_EDEVICE, "internal error: more BFS levels than vertices");
                batch = std::min(batch * 2, 64);
                enqueue(batch, b & 1);
                if (done((b - 1) & 1)) break;
            }
        }
    }
    GG_HIP(hipEventRecord(ev_stop, stream));
    unsigned long long fin[kCtlWords];
    GG_HIP(hipMemcpyAsync(fin, ws.ctl, sizeof(fin), hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
    if (fin[CTL_ERR] & ERR_TILE) fail(GB_EDEVICE, "internal error: inconsistent frontier tile index");
    if (fin[CTL_ERR] & ERR_RANGE) fail(GB_EDEVICE, "internal error: frontier entry or column index out of range");
    if (!fin[CTL_DONE]) fail(GB_EDEVICE, "internal error: BFS did not reach an empty frontier");
    last_num_levels = static_cast<int>(fin[CTL_MAXLEV]) + 1;
    float ms = 0.f;
    GG_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
    last_ms = ms;
    if (st) {
        float kms = 0.f;
        GG_HIP(hipEventElapsedTime(&kms, ev_kernel, ev_stop));
        st->ms = ms;
        st->kernel_ms = kms;
        st->num_levels = last_num_levels;
        st->reserved = 0;
        st->n_reached = static_cast<int64_t>(fin[CTL_NREACH]);
        st->m_reached = static_cast<int64_t>(fin[CTL_MREACH]);
    }
}

void Runner::bfs_to_host(int32_t source, int32_t* h_levels, int32_t* h_parents, gb_stats* st) {
    if (!h_levels) fail(GB_EINVAL, "levels buffer is null");
    DeviceGuard guard(device);
    ensure_workspace();
    bfs(source, ws.levels, h_parents ? ws.parents : nullptr, st);
    GG_HIP(hipMemcpyAsync(h_levels, ws.levels, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
    if (h_parents) GG_HIP(hipMemcpyAsync(h_parents, ws.parents, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
}

}  // namespace gg
