// Single-source BFS for gfx950, direction-optimising (top-down / bottom-up).
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
//  * A level L runs as `expand(L)` (top-down or bottom-up, chosen on the
//    device) and `compact(L)`, issued back to back with no host
//    synchronisation; the host only checks a done flag once per batch of
//    levels.
//  * Top-down: the frontier of a level is a queue of entries (vertex, first
//    CSR entry, exclusive prefix of degrees).  expand(L) cuts the level's
//    edges into tiles of T = BLOCK*K edges; tile_start[t] names the entry
//    owning edge t*T, so each workgroup loads its entries with one lookup and
//    every thread owns K edges of the tile (coalesced column reads) whatever
//    the degree distribution: hubs are spread over many workgroups.
//  * Claims: plain load of the visited-bitmap word (may be stale, never
//    wrongly "visited"), then atomicOr; the winner writes level L+1 and the
//    parent.  Levels are deterministic.
//  * Sparse levels (at most `dense_edges` frontier edges) build the next
//    queue directly: winners go to an LDS list (wave-aggregated LDS atomics,
//    ballot + mbcnt), get a block scan, and ONE 64-bit atomicAdd per tile on
//    a packed (count << 32 | edges) counter gives both their queue slots and
//    their edge offsets.
//  * Dense levels skip the enqueue.  compact(L) rebuilds the queue from the
//    bitmaps: new = visited & ~prev over 16384-vertex chunks -- per-thread
//    serial sums, a wave64 scan and an LDS-staged block scan give each new
//    vertex its offset inside the chunk, and the chunk reserves its queue
//    slots and edge range with one packed 64-bit atomicAdd, like a sparse
//    tile.  It also stores the new vertices as a frontier bitmap.
//  * Bottom-up (Beamer, Asanovic, Patterson, SC 2012), on symmetric graphs:
//    when a level's frontier edges exceed the unexplored edges / alpha, each
//    unvisited vertex looks for a neighbour in the frontier bitmap and stops
//    at the first one.  A wave owns groups of 64 consecutive vertices (two
//    bitmap words each, so it updates them with plain stores).  Each lane
//    probes the first entries of its own row; rows still open after those go
//    to 8-lane sub-waves that scan 8 entries per step -- the degree tiers of
//    a thread or a group of lanes per vertex.  Isolated vertices are marked
//    visited up front.  Search goes bottom-up only while the frontier grows,
//    and back top-down once it has fewer than n / beta vertices and shrinks
//    (Beamer's rules); compact() then turns the frontier bitmap into a queue.
//  * Directions are chosen on the device from counters that no kernel of
//    the level modifies, so every workgroup of expand(L) and compact(L)
//    agrees.  The host allows bottom-up (and keeps the frontier bitmaps it
//    needs) only for levels where searches of this graph have gone
//    bottom-up (a window learned per handle); elsewhere the level runs
//    top-down, which is always correct.
//  * No workgroup waits on another, so the kernels need no particular
//    residency and can share the device with other streams' work.
//  * Every kernel stops at once when the error word is set, and the tiles
//    check each queue entry and column index against n and m, so an
//    inconsistent queue is reported as an error, never dereferenced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device_graph.hpp"
#include "frontier.hpp"
#include "wave.hpp"

namespace gg {

namespace {

// ctl[] layout (64-bit words).
enum : int {
    CTL_RING0 = 0,     // 3 words: (count << 32 | edges) of level L's frontier at ring[L % 3]
    CTL_NREACH = 3,    // reached vertices
    CTL_MREACH = 4,    // CSR entries of the frontiers expanded so far (compact(L) adds level L's)
    CTL_MAXLEV = 5,    // largest level assigned
    CTL_DONE = 6,      // set by the first expand() that finds an empty frontier
    CTL_ERR = 7,       // internal consistency error (bit set)
    CTL_REP0 = 8,      // 3 words: forms of level L's frontier at rep[L % 3] (REP_* bits)
    CTL_DIR0 = 11,     // 3 words: direction of level L at dir[L % 3]
    CTL_QRES = 14,     // reservations when a frontier bitmap is turned into a queue
    CTL_MAXE = 15,     // largest frontier seen (CSR entries)
    CTL_BUFIRST = 16,  // first / last level searched bottom-up (~0 / 0 if none)
    CTL_BULAST = 17,
    CTL_EXP0 = 18,     // 2 words: CSR entries of the frontiers of levels < L at exp[L & 1]
    CTL_DCNT0 = 20,    // 2 words: (count << 32 | edges) of the winners of a dense level L at dcnt[(L+1) & 1]
    CTL_FCNT0 = 22,    // 2 words, one-launch levels: (count << 32 | edges) of level L at fcnt[L & 1]
    CTL_RCNT0 = 24,    // 2 words, one-launch levels: queued vertices of levels <= L at rcnt[L & 1]
    CTL_DEEPE0 = 26,   // 2 words, deep levels: CSR entries of level L's frontier at deepe[L & 1]
    CTL_TAIL = 28,     // what the tail kernel did: GB_TAIL_* outcome | (level << 8)
    CTL_TAILFROM = 29, // first level of the current run of small queued levels (~0: none)
    CTL_T0 = 30,       // device clock (wall_clock64) when the clear kernel started
    CTL_T1 = 31,       // in a stored snapshot only: device clock when it was taken
    CTL_WORDS = 32
};
// What the host fills a snapshot slot with before the launch: no control word
// takes this value (counts and packed counts stay below 2^63, flags and
// levels are small, "none" is ~0, the clock would need 10^11 s to get there).
constexpr unsigned long long kSnapUnwritten = 0xA5A5A5A55A5A5A5Aull;
enum : unsigned long long { ERR_TILE = 1, ERR_RANGE = 2 };
enum : unsigned long long { REP_QUEUE = 1, REP_BITMAP = 2 };
enum : unsigned long long { DIR_TD = 1, DIR_BU = 2 };
// BfsParams.flags of a direct launch.
enum : int32_t {
    BFS_PARAM_INLINE_L0 = 1,  // the clear kernel runs level 0 (the host starts at level 1)
    BFS_PARAM_BU_NEXT = 2,    // ... and level 1 may go bottom-up: its winners also go to fb[1]
    BFS_PARAM_ROWS_UNIQUE = 4,  // ... and no row repeats a neighbour (no duplicate check)
};
constexpr int kClearBlock = 256;
constexpr uint32_t kInlineL0 = 128;  // largest source degree whose level 0 the clear kernel runs

constexpr int kCompactBlock = 1024;
constexpr int kCompactRounds = 16;                          // vertices per thread
constexpr int kChunk = kCompactBlock * kCompactRounds;      // vertices per compaction chunk
constexpr int kBuGroups = 4;                                // 64-vertex groups per wave pass
constexpr int kBuFirst = 4;                                 // row entries each lane probes before the list (= int4 head)
static_assert(kBuFirst == 4, "the head array holds one int4 per vertex");
constexpr int kMaxCapturedLevels = 32;                      // deeper searches continue with direct launches
constexpr int kDeepDepth = 64;                              // learned depth from which GB_DEEP_AUTO applies
constexpr int kTailBlock = 1024;                            // the tail kernel's one workgroup
constexpr int kTailCap = 1536;                              // frontier entries of a tail level (LDS)
constexpr uint32_t kTailMaxEdges = 1u << 15;                // CSR entries of a level the tail expands
constexpr uint32_t kTailLearnEdges = 4096;                  // levels this small count toward the tail start
constexpr int kTailK = 4;                                   // edges per thread per pass
constexpr int kNoTail = 1 << 20;                            // learned tail start when the last level was large
constexpr uint32_t kGrowDenseEdges = 4096;                  // dense threshold of the levels before the first bottom-up one

struct BfsArgs {
    const int64_t* __restrict__ offsets;
    const int32_t* __restrict__ cols;
    uint32_t* vis;    // claimed vertices
    uint32_t* prev;   // vertices already placed in (or handled by) a frontier queue
    uint32_t* fb[3];  // frontier bitmaps: level L's at fb[L % 3]
    const uint32_t* iso;  // vertices of degree 0 (symmetric graphs), or null
    const int4* head;     // first kBuFirst neighbours of each vertex, -1 padded, the 4th u stored as -2 - u
                          // when the row is longer (symmetric graphs)
    int64_t* q_rb[2];
    uint32_t* q_eoff[2];
    int32_t* q_v[2];
    uint32_t* tile_start[2];
    unsigned long long* ctl;
    BfsParams* p;               // source, levels, parents of this call (device memory)
    const BfsParams* hp;        // the same, host-visible (written by the host before the launch)
    int64_t n;
    int64_t m;
    uint32_t nwords;
    uint32_t nchunks;
    uint32_t dense_edges;
    uint32_t tile_edges;        // T of the expand kernel (tile_start granularity)
    uint32_t bottom_up;         // direction optimisation enabled
    uint32_t fused;             // one launch per level (symmetric graphs, not deep mode)
    int64_t n_edged;            // vertices with at least one CSR entry
    uint32_t alpha, beta;
    int32_t grow_until;         // levels L with L+1 <= grow_until count as dense from kGrowDenseEdges
};

// Beamer's test for a frontier of E CSR entries after `explored` entries of
// earlier frontiers: frontier edges * alpha against the unexplored edges.
// Top-down -> bottom-up (Beamer et al.): the frontier's edges E, growing
// from the previous level's E_prev, exceed the unexplored edges / alpha.
// (Unexplored edges include those of vertices the search never reaches, so
// without the growth test a small frontier late in the search would go
// bottom-up.)
__device__ __forceinline__ bool go_bottom_up(const BfsArgs& a, int64_t explored, uint32_t E, uint32_t E_prev) {
    const int64_t unexplored = a.m - explored - static_cast<int64_t>(E);
    return E > E_prev && static_cast<int64_t>(E) * a.alpha > unexplored;
}

// Bottom-up -> top-down (Beamer et al.): the next frontier has fewer than
// n / beta vertices and is shrinking.
__device__ __forceinline__ bool go_top_down(const BfsArgs& a, uint32_t F, uint32_t F_next) {
    return F_next < F && static_cast<int64_t>(F_next) * a.beta < a.n;
}

// Direction of level L.  Reads only words that no kernel of level L writes.
// F, E: level L's vertices and CSR entries; F_prev, E_prev_fused: level
// L-1's (one-launch levels only).
__device__ __forceinline__ unsigned long long choose_dir(const BfsArgs& a, int L, uint32_t F, uint32_t E,
                                                         uint32_t F_prev, uint32_t E_prev_fused, bool bu_launched) {
    if (!a.bottom_up || !bu_launched || L == 0) return DIR_TD;
    const unsigned long long rep = a.ctl[CTL_REP0 + L % 3];
    if (!(rep & REP_BITMAP)) return DIR_TD;
    // Only the bitmap: a bottom-up level that stays bottom-up, or (one-launch
    // levels) a dense top-down level.
    if (!(rep & REP_QUEUE)) return DIR_BU;
    if (a.ctl[CTL_DIR0 + (L - 1) % 3] == DIR_BU) {
        // compact(L-1) queues level L only to go back top-down; a one-launch
        // bottom-up level queues it ahead when its frontier shrinks, and the
        // Beamer test is made here.
        if (!a.fused) return DIR_TD;
        return go_top_down(a, F_prev, F) ? DIR_TD : DIR_BU;
    }
    // After a top-down level: frontier edges against the unexplored edges.
    // EXP[L & 1] = edges of levels <= L-1, EXP[(L + 1) & 1] = of levels <= L-2;
    // compact(L) is the next kernel to write either.  A one-launch level
    // writes EXP[(L + 1) & 1] itself (block 0, while the others decide), so
    // it takes level L-1's edges from fcnt instead.
    const unsigned long long through_prev = a.ctl[CTL_EXP0 + (L & 1)];
    const uint32_t E_prev =
        a.fused ? E_prev_fused : static_cast<uint32_t>(through_prev - a.ctl[CTL_EXP0 + ((L + 1) & 1)]);
    return go_bottom_up(a, static_cast<int64_t>(through_prev), E, E_prev) ? DIR_BU : DIR_TD;
}

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
__device__ void expand_tile(const BfsArgs& a, int32_t* levels, int32_t* parents, TileLds<BLOCK, K>& s, int L,
                            uint32_t t, uint32_t ntiles, uint32_t F, uint32_t E, bool dense, bool bitmap_next,
                            uint32_t& dn_cnt, uint64_t& dn_edges, uint32_t& dn_zero) {
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
            levels[uu] = L + 1;
            if (parents) parents[uu] = s.v[owner[k]];
        }
    }
    if (dense) {
        // compact() builds the next queue (or bitmap) from the bitmaps; it
        // needs the next level's totals first, so the winners are counted here.
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (won[k]) {
                const uint32_t uu = static_cast<uint32_t>(u[k]);
                // Without compact() the winners mark level L+1's bitmap here.
                if (bitmap_next && a.fused) atomicOr(&a.fb[(L + 1) % 3][uu >> 5], 1u << (uu & 31u));
                const uint32_t d = static_cast<uint32_t>(a.offsets[uu + 1] - a.offsets[uu]);
                if (d) {
                    ++dn_cnt;
                    dn_edges += d;
                } else {
                    ++dn_zero;
                }
            }
        }
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
            if (bitmap_next) atomicOr(&a.fb[(L + 1) % 3][uu >> 5], 1u << (uu & 31u));
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

// Level L bottom-up (run by the level kernel when level L goes bottom-up):
// every unvisited vertex looks for a neighbour in the frontier bitmap of
// level L; the vertices found form the bitmap of level L+1, and their
// (count, CSR entries) go to ring[(L+1) % 3].
//
// The step is bound by memory latency, not bandwidth (a search level
// touches few bytes per vertex), so it is organised for loads in flight:
//   * a wave owns chunks of kBuChunkGroups groups of 64 vertices (their
//     visited words, one load per lane) and lists the chunk's unvisited
//     vertices in LDS first: a late level, where most vertices are visited,
//     then costs one round per chunk instead of one pass per 4 groups;
//   * a round takes kBuGroups list entries per lane and issues the
//     first-entries, offset and frontier-probe loads of all of them
//     together (two dependent steps per round);
//   * every listed lane probes the first kBuFirst neighbours of its vertex,
//     kept side by side per vertex (`head`, built at upload); the first hit
//     usually comes within them;
//   * rows not settled by that go to a per-wave list in LDS, which 8-lane
//     sub-waves drain once per chunk (or when it fills up), not once per
//     round: a sub-wave scans 8 consecutive entries of its row per step,
//     stops at the first hit, and takes the next row from the list as soon
//     as it is done, so long and short rows share the wave's steps.
// The vertices found are collected as bits in LDS; the wave owns its
// chunk's bitmap words and writes them with plain stores.  The output
// bitmap arrives zeroed (two levels ahead), so only words with new vertices
// are written.
constexpr int kBuChunkGroups = 16;                 // groups of 64 vertices per wave chunk
constexpr int kBuChunk = kBuChunkGroups * 64;      // vertices per wave chunk
constexpr int kBuChunkWords = kBuChunkGroups * 2;  // their visited words (lanes 0..31)
constexpr int kBuList = kBuGroups * 64 + 128;      // long rows listed before the wave drains them
static_assert(kBuChunkWords <= 64, "one visited word per lane");
static_assert(kBuList >= kBuGroups * 64, "a round adds at most kBuGroups * 64 long rows");
struct BuWaveLds {
    int64_t rb[kBuList];            // long rows: next entry of the row
    uint32_t len[kBuList];          // long rows: entries left
    uint16_t vid[kBuList];          // long rows: the vertex (offset in the chunk)
    uint16_t unv[kBuChunk];         // the chunk's unvisited vertices (offsets in the chunk)
    uint32_t found[kBuChunkWords];  // the chunk's vertices found at this level
};

template <int WAVES>
__device__ void bottom_up_level(const BfsArgs& a, int L, BuWaveLds* s_wave, bool enqueue) {
    __shared__ unsigned long long s_acc[WAVES];
    BuWaveLds& sw = s_wave[threadIdx.x >> 6];
    const uint32_t* fin = a.fb[L % 3];
    uint32_t* fout = a.fb[(L + 1) % 3];
    int32_t* const levels = a.p->levels;
    int32_t* const parents = a.p->parents;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t sg = lane >> 3, sl = lane & 7u;  // sub-wave and lane within it
    const uint64_t n = static_cast<uint64_t>(a.n);
    const uint64_t nchunks = (n + kBuChunk - 1) / kBuChunk;
    const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * WAVES;
    uint32_t my_cnt = 0;
    uint64_t my_edges = 0;
    bool bad = false;
    auto in_frontier = [&](int32_t u) -> bool {
        if (static_cast<uint32_t>(u) >= n) {
            bad = true;
            return false;
        }
        return (fin[static_cast<uint32_t>(u) >> 5] >> (u & 31)) & 1u;
    };
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * WAVES + (threadIdx.x >> 6); c < nchunks; c += nwaves) {
        // The chunk's visited words (lanes 0..31) and its unvisited vertices,
        // listed in LDS in vertex order.
        const uint64_t w = c * kBuChunkWords + lane;
        uint32_t vw = ~0u;
        if (lane < static_cast<uint32_t>(kBuChunkWords) && w < a.nwords) {
            vw = a.vis[w];
            const uint64_t v0 = w * 32;  // bits of vertices >= n count as visited
            if (v0 + 32 > n) vw |= ~0u << static_cast<uint32_t>(n - v0);
        }
        const uint32_t mine = static_cast<uint32_t>(__popc(~vw));
        const uint32_t incl = wave_inclusive_scan(mine);
        const uint32_t U = __shfl(incl, 63, 64);
        if (U == 0) continue;  // wave-uniform
        {
            uint32_t pos = incl - mine;
            for (uint32_t b = ~vw; b; b &= b - 1) sw.unv[pos++] = static_cast<uint16_t>(lane * 32 + __ffs(b) - 1);
        }
        if (lane < static_cast<uint32_t>(kBuChunkWords)) sw.found[lane] = 0u;
        wave_lds_sync();
        // Rows not settled by their first entries wait in the wave's list and
        // are drained together (once per chunk, unless the list fills up):
        // 8 rows at a time, one per 8-lane sub-wave, 8 entries per step; a
        // sub-wave stops at its row's first hit and takes the next row at
        // once.  The drain stores the found vertices' levels itself.
        uint32_t nlist = 0;
        auto drain = [&]() {  // wave-uniform
            wave_lds_sync();
            int32_t item = sg < nlist ? static_cast<int32_t>(sg) : -1;
            int64_t pos = 0, end = 0;
            if (item >= 0) {
                pos = sw.rb[item];
                end = pos + sw.len[item];
            }
            uint32_t next = 8;
            while (wave_ballot(item >= 0)) {
                const int64_t i = pos + sl;
                int32_t cu = -1;
                bool hit = false;
                if (item >= 0 && i < end) {
                    cu = a.cols[i];
                    hit = in_frontier(cu);
                }
                const uint64_t hb = wave_ballot(hit);
                const uint32_t mine8 = static_cast<uint32_t>(hb >> (sg * 8)) & 0xffu;
                const int src = mine8 ? static_cast<int>(sg * 8 + __ffs(mine8) - 1) : static_cast<int>(lane);
                const int32_t pu = __shfl(cu, src, 64);
                bool done = false;
                if (item >= 0) {
                    if (mine8) {
                        if (sl == 0) {
                            const uint32_t off = sw.vid[item];
                            const int64_t v = static_cast<int64_t>(c * kBuChunk + off);
                            levels[v] = L + 1;
                            if (parents) parents[v] = pu;
                            atomicOr(&sw.found[off >> 5], 1u << (off & 31u));
                            if (!enqueue) {  // with `enqueue` the block counts them below
                                ++my_cnt;
                                my_edges += static_cast<uint64_t>(sw.len[item]) + kBuFirst;
                            }
                        }
                        done = true;
                    } else {
                        pos += 8;
                        done = pos >= end;
                    }
                }
                const uint64_t need = wave_ballot(done && sl == 0);
                if (done) {
                    const uint32_t t = next + static_cast<uint32_t>(__popcll(need & ((1ull << (sg * 8)) - 1)));
                    item = t < nlist ? static_cast<int32_t>(t) : -1;
                    if (item >= 0) {
                        pos = sw.rb[item];
                        end = pos + sw.len[item];
                    }
                }
                next += static_cast<uint32_t>(__popcll(need));
            }
            wave_lds_sync();  // the list is refilled from slot 0
            nlist = 0;
        };
        for (uint32_t r0 = 0; r0 < U; r0 += kBuGroups * 64) {
            if (nlist > static_cast<uint32_t>(kBuList - kBuGroups * 64)) drain();
            int64_t vtx[kBuGroups];
            int64_t rb[kBuGroups], re[kBuGroups];
            int32_t par[kBuGroups];
            // Step 1: first neighbours (side by side per vertex) and row
            // bounds of the listed vertices.
            int4 hd[kBuGroups];
#pragma unroll
            for (int j = 0; j < kBuGroups; ++j) {
                const uint32_t i = r0 + j * 64 + lane;
                vtx[j] = i < U ? static_cast<int64_t>(c * kBuChunk + sw.unv[i]) : -1;
                rb[j] = re[j] = 0;
                hd[j] = make_int4(-1, -1, -1, -1);
                if (vtx[j] >= 0) {
                    hd[j] = a.head[vtx[j]];
                    rb[j] = a.offsets[vtx[j]];
                    re[j] = a.offsets[vtx[j] + 1];
                }
            }
            // Step 2: probes of the first neighbours; rows still open go to
            // the long-row list, the others are settled here.
#pragma unroll
            for (int j = 0; j < kBuGroups; ++j) {
                const int32_t u[kBuFirst] = {hd[j].x, hd[j].y, hd[j].z, hd[j].w < -1 ? -2 - hd[j].w : hd[j].w};
                par[j] = -1;
#pragma unroll
                for (int k = 0; k < kBuFirst; ++k)
                    if (par[j] < 0 && u[k] >= 0 && in_frontier(u[k])) par[j] = u[k];
                const bool cc = vtx[j] >= 0 && par[j] < 0 && re[j] - rb[j] > kBuFirst;
                const uint64_t m = wave_ballot(cc);
                if (cc) {
                    const uint32_t slot = nlist + rank_below(m);
                    sw.rb[slot] = rb[j] + kBuFirst;
                    sw.len[slot] = static_cast<uint32_t>(re[j] - rb[j] - kBuFirst);
                    sw.vid[slot] = static_cast<uint16_t>(vtx[j] - static_cast<int64_t>(c * kBuChunk));
                }
                nlist += static_cast<uint32_t>(__popcll(m));
                if (vtx[j] >= 0 && par[j] >= 0) {
                    const int64_t v = vtx[j];
                    levels[v] = L + 1;
                    if (parents) parents[v] = par[j];
                    const uint32_t off = static_cast<uint32_t>(v - static_cast<int64_t>(c * kBuChunk));
                    atomicOr(&sw.found[off >> 5], 1u << (off & 31u));
                    if (!enqueue) {  // with `enqueue` the block counts them below
                        ++my_cnt;
                        my_edges += static_cast<uint64_t>(re[j] - rb[j]);
                    }
                }
            }
        }
        if (nlist) drain();
        wave_lds_sync();
        if (lane < static_cast<uint32_t>(kBuChunkWords) && w < a.nwords) {
            const uint32_t bits = sw.found[lane];
            if (bits) {  // fout was zeroed two levels ahead
                fout[w] = bits;
                a.vis[w] = vw | bits;
                a.prev[w] = vw | bits;
            }
        }
        wave_lds_sync();  // the next chunk's list reuses the LDS
    }
    if (enqueue) {
        // Queue level L+1 (a top-down level may follow) from the bitmap words
        // this workgroup wrote: in each pass its waves covered WAVES
        // consecutive chunks (kBuChunkWords words each).  The words of all its
        // passes are taken at once, spread over the workgroup's threads, and
        // one packed (count << 32 | edges) reservation covers them all, as in
        // compact(); a winner has an edge (to its parent).  The barrier makes
        // the waves' word stores visible.
        __syncthreads();
        constexpr uint32_t kBlock = WAVES * 64;
        constexpr uint32_t kWords = WAVES * kBuChunkWords;  // words per pass
        __shared__ unsigned long long s_wave[WAVES];
        __shared__ unsigned long long s_base;
        const int q = (L + 1) & 1;
        const uint64_t first = static_cast<uint64_t>(blockIdx.x) * WAVES;  // first chunk of each pass
        const uint64_t pstride = static_cast<uint64_t>(gridDim.x) * WAVES;
        const uint64_t npass = first < nchunks ? (nchunks - first + pstride - 1) / pstride : 0;
        const uint64_t nslots = npass * kWords;
        // Word slot j: word j % kWords of pass j / kWords.
        auto word_of = [&](uint64_t j) { return (first + (j / kWords) * pstride) * kBuChunkWords + j % kWords; };
        uint32_t cnt = 0, edges = 0;
        for (uint64_t j = threadIdx.x; j < nslots; j += kBlock) {
            const uint64_t w = word_of(j);
            const uint32_t bits = w < a.nwords ? fout[w] : 0u;
            for (uint32_t b = bits; b; b &= b - 1) {
                const int64_t v = static_cast<int64_t>(w * 32 + __ffs(b) - 1);
                ++cnt;
                edges += static_cast<uint32_t>(a.offsets[v + 1] - a.offsets[v]);
            }
        }
        const unsigned long long mine = (static_cast<unsigned long long>(cnt) << 32) | edges;
        unsigned long long total;
        const unsigned long long excl = block_exclusive_scan<kBlock>(mine, s_wave, total);
        if (threadIdx.x == 0) {
            s_base = total ? atomicAdd(&a.ctl[CTL_RING0 + (L + 1) % 3], total) : 0ull;
            if (total && !reservation_fits(s_base, total, a.n, a.m)) {
                atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
                s_base = ~0ull;
            }
        }
        __syncthreads();
        const unsigned long long base = s_base;
        if (cnt && base != ~0ull) {
            uint32_t slot = static_cast<uint32_t>((base + excl) >> 32);
            uint32_t eo = static_cast<uint32_t>(base + excl);
            for (uint64_t j = threadIdx.x; j < nslots; j += kBlock) {
                const uint64_t w = word_of(j);
                const uint32_t bits = w < a.nwords ? fout[w] : 0u;
                for (uint32_t b = bits; b; b &= b - 1) {
                    const int64_t v = static_cast<int64_t>(w * 32 + __ffs(b) - 1);
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
        __syncthreads();  // s_wave is reused below
    }
    if (bad) atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
    const unsigned long long mine = (static_cast<unsigned long long>(wave_sum(my_cnt)) << 32) | wave_sum(my_edges);
    if (lane == 0) s_acc[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < WAVES; ++i) tot += s_acc[i];
        if (tot) atomicAdd(&a.ctl[CTL_RING0 + (L + 1) % 3], tot);
    }
}

// Level L, step 1: pick the direction, then expand level L top-down (the
// queue) or bottom-up (the bitmap).  `bu_launched`: level L may go
// bottom-up (it lies in the learned window, and the bitmaps it needs are
// kept); `bu_next`: level L+1 may, so a sparse level also records its
// winners in the frontier bitmap fb[(L+1) & 1] (zeroed by compact(L-1)).
template <int BLOCK, int K>
union LevelLds {
    TileLds<BLOCK, K> td;
    BuWaveLds bu[BLOCK / 64];
};

template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void bfs_expand_kernel(BfsArgs a, int L, int bu_launched, int bu_next, int deep) {
    __shared__ LevelLds<BLOCK, K> lds;
    TileLds<BLOCK, K>& s = lds.td;
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
    // Level L-1's (count << 32 | edges), kept by one-launch levels.
    const unsigned long long fc_prev = a.ctl[CTL_FCNT0 + ((L + 1) & 1)];
    const uint32_t F_prev = static_cast<uint32_t>(fc_prev >> 32), E_prev = static_cast<uint32_t>(fc_prev);
    const unsigned long long dir = choose_dir(a, L, F, E, F_prev, E_prev, bu_launched != 0);
    // A dense level leaves level L+1 in the bitmaps only.  Without compact()
    // (one-launch levels) that is allowed only if level L+1 may go bottom-up,
    // which it then must.  Up to the first level that searches of this graph
    // took bottom-up (grow_until), levels grow fast enough that a smaller
    // frontier already counts as dense: the next level goes bottom-up anyway,
    // and skipping its queue shortens this level's chain.
    const uint32_t dense_min = (a.fused && L + 1 <= a.grow_until) ? kGrowDenseEdges : a.dense_edges;
    const bool dense = !deep && E > dense_min && (!a.fused || bu_next);
    // A one-launch bottom-up level also queues level L+1 when that level
    // cannot go bottom-up, or may go back top-down: the frontier shrinks, or
    // so few vertices are left that level L+1 is below n / beta whatever it
    // holds.
    const unsigned long long through = a.ctl[CTL_RCNT0 + ((L + 1) & 1)] + F;  // queued, levels <= L
    const bool few_left = (a.n_edged - static_cast<int64_t>(through)) * a.beta < a.n;
    const bool bu_queue = a.fused && dir == DIR_BU && (!bu_next || F < F_prev || few_left);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // ring[(L+2)%3] held level L-1's counter; it next counts level L+2.
        a.ctl[CTL_RING0 + (L + 2) % 3] = 0;
        atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(F));
        atomicMax(&a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L));
        atomicMax(&a.ctl[CTL_MAXE], static_cast<unsigned long long>(E));
        a.ctl[CTL_DIR0 + L % 3] = dir;
        a.ctl[CTL_REP0 + (L + 1) % 3] =
            dir == DIR_BU ? (bu_queue ? REP_QUEUE | REP_BITMAP : REP_BITMAP)
                          : (dense ? (a.fused ? REP_BITMAP : 0ull) : (bu_next ? REP_QUEUE | REP_BITMAP : REP_QUEUE));
        a.ctl[CTL_QRES] = 0;
        if (deep || a.fused) {
            // One-launch levels have no compact(L): its bookkeeping is done
            // here (the search ends when expand(L+1) finds level L+1 empty).
            atomicAdd(&a.ctl[CTL_MREACH], static_cast<unsigned long long>(E));
            a.ctl[CTL_EXP0 + ((L + 1) & 1)] = a.ctl[CTL_EXP0 + (L & 1)] + E;
        }
        if (a.fused) {
            a.ctl[CTL_FCNT0 + (L & 1)] = cnt;
            a.ctl[CTL_RCNT0 + (L & 1)] = through;
            // Where the search's tail starts (the host learns it): the first
            // level of the last run of small levels that have a queue.
            const bool small = E <= kTailLearnEdges && F <= kTailCap && (a.ctl[CTL_REP0 + L % 3] & REP_QUEUE);
            if (!small) a.ctl[CTL_TAILFROM] = ~0ull;
            else if (a.ctl[CTL_TAILFROM] == ~0ull) a.ctl[CTL_TAILFROM] = static_cast<unsigned long long>(L);
        }
    }
    if (a.fused) {
        // fb[(L+2) % 3] held level L-1's frontier; level L+1 writes its
        // winners there.
        uint32_t* fz = a.fb[(L + 2) % 3];
        for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < a.nwords; i += gridDim.x * BLOCK) fz[i] = 0u;
    }
    if (dir == DIR_BU) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicMin(&a.ctl[CTL_BUFIRST], static_cast<unsigned long long>(L));
            atomicMax(&a.ctl[CTL_BULAST], static_cast<unsigned long long>(L));
        }
        bottom_up_level<BLOCK / 64>(a, L, lds.bu, bu_queue);
        return;
    }
    int32_t* const levels = a.p->levels;
    int32_t* const parents = a.p->parents;
    const uint32_t ntiles = static_cast<uint32_t>((uint64_t(E) + T - 1) / T);
    uint32_t dn_cnt = 0, dn_zero = 0;
    uint64_t dn_edges = 0;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x)
        expand_tile<BLOCK, K>(a, levels, parents, s, L, t, ntiles, F, E, dense, bu_next != 0, dn_cnt, dn_edges,
                              dn_zero);
    if (deep) return;  // level L is sparse: its queue is built
    if (dense) {
        // One atomic per workgroup: the next level's (count, edges) and the
        // degree-0 winners (reached, never queued).
        const unsigned long long mine =
            (static_cast<unsigned long long>(wave_sum(dn_cnt)) << 32) | wave_sum(dn_edges);
        const uint32_t z = wave_sum(dn_zero);
        __shared__ unsigned long long s_tot[BLOCK / 64];
        __shared__ uint32_t s_z[BLOCK / 64];
        if ((threadIdx.x & 63) == 0) {
            s_tot[threadIdx.x >> 6] = mine;
            s_z[threadIdx.x >> 6] = z;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long tot = 0;
            uint32_t zt = 0;
            for (int i = 0; i < BLOCK / 64; ++i) {
                tot += s_tot[i];
                zt += s_z[i];
            }
            // compact(L) reads the count (and builds the queue); one-launch
            // levels count level L+1 where expand(L+1) reads it.
            if (tot) atomicAdd(&a.ctl[a.fused ? CTL_RING0 + (L + 1) % 3 : CTL_DCNT0 + ((L + 1) & 1)], tot);
            if (zt) {
                atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(zt));
                atomicMax(&a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L + 1));
            }
        }
    }
}

// Level L, step 3: queue of level L+1 from a bitmap, when level L+1 runs
// top-down and its queue was not built during level L.  Chunk c covers
// vertices [c*kChunk, (c+1)*kChunk); thread t takes vertices
// c*kChunk + r*kCompactBlock + t for r < 16, so the offsets loads of a wave are
// contiguous.  The thread's new vertices occupy consecutive queue slots; the
// chunk's block goes where its atomicAdd on the counter puts it.
//   from_bitmap = false (dense top-down level): new = vis & ~prev; also
//       stores the new vertices in fb[(L+1) & 1] and sets prev = vis;
//       reserves on ring[(L+1) % 3].
//   from_bitmap = true (after a bottom-up level): new = fb[(L+1) & 1];
//       ring[(L+1) % 3] already holds the totals, so it reserves on QRES.
__device__ void compact_chunk(const BfsArgs& a, int L, uint32_t c, bool from_bitmap) {
    __shared__ unsigned long long s_wave[kCompactBlock / 64];
    __shared__ unsigned long long s_excl;
    __shared__ uint32_t s_ok;
    const uint32_t tid = threadIdx.x;
    const int64_t vbase = int64_t(c) * kChunk;
    const uint32_t* fnew = a.fb[(L + 1) % 3];
    uint32_t newmask = 0;  // bit r: vertex vbase + r*kCompactBlock + tid is new (degree > 0)
    uint32_t kept = 0, edges = 0;
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        const int64_t v = vbase + r * kCompactBlock + tid;
        if (v < a.n) {
            const uint32_t w = static_cast<uint32_t>(v >> 5);
            const uint32_t nb = from_bitmap ? fnew[w] : (a.vis[w] & ~a.prev[w]);
            if ((nb >> (v & 31)) & 1u) {
                const uint32_t d = static_cast<uint32_t>(a.offsets[v + 1] - a.offsets[v]);
                if (d) {
                    newmask |= 1u << r;
                    ++kept;
                    edges += d;
                }
            }
        }
    }
    // Block scan of (kept, edges) packed in 64 bits.  Degree-0 new vertices
    // were counted when they were claimed.
    const unsigned long long mine = (static_cast<unsigned long long>(kept) << 32) | edges;
    unsigned long long total;
    const unsigned long long excl_local = block_exclusive_scan<kCompactBlock>(mine, s_wave, total);
    if (!from_bitmap) {
        // fb[(L+1)&1] = new, prev |= vis for the chunk's words: every read of
        // prev above happened before the block scan's barriers.
        uint32_t* fout = a.fb[(L + 1) % 3];
        for (uint32_t i = tid; i < kChunk / 32; i += kCompactBlock) {
            const int64_t w = vbase / 32 + i;
            if (w < a.nwords) {
                const uint32_t vw = a.vis[w];
                fout[w] = vw & ~a.prev[w];
                a.prev[w] = vw;
            }
        }
    }
    if (tid == 0) {
        s_ok = 1;
        if (total) {
            s_excl = atomicAdd(&a.ctl[from_bitmap ? CTL_QRES : CTL_RING0 + (L + 1) % 3], total);
            if (!reservation_fits(s_excl, total, a.n, a.m)) {
                atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
                s_ok = 0;
            }
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

// `bu_next`: level L+1 may go bottom-up; `bu_next2`: level L+2 may.  With
// either, fb[(L+2) % 3] (level L-1's frontier, next written as level L+2's
// bitmap by a bottom-up or sparse top-down expand(L+1), which only set bits)
// is zeroed here.  Level L+1's
// counter ring[(L+1)%3] was zeroed by expand(L-1); a dense top-down level L
// left it alone and the chunks' reservations build it up.
__global__ __launch_bounds__(kCompactBlock) void bfs_compact_kernel(BfsArgs a, int L, int bu_next, int bu_next2) {
    if (a.ctl[CTL_DONE] | a.ctl[CTL_ERR]) return;
    if (bu_next || bu_next2) {
        // fb[L & 1] is the output bitmap of level L+1: bottom-up and sparse
        // top-down winners only set bits in it.
        uint32_t* fz = a.fb[(L + 2) % 3];
        for (uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x; i < a.nwords; i += gridDim.x * kCompactBlock)
            fz[i] = 0u;
    }
    const unsigned long long cnt = a.ctl[CTL_RING0 + L % 3];
    const uint32_t F = static_cast<uint32_t>(cnt >> 32), E = static_cast<uint32_t>(cnt);
    if (F == 0) return;
    const int64_t explored = static_cast<int64_t>(a.ctl[CTL_EXP0 + (L & 1)]) + E;  // levels <= L
    const unsigned long long dir = a.ctl[CTL_DIR0 + L % 3];
    // Level L+1's vertex count is final here on every path: bottom-up and
    // sparse top-down levels have filled ring[(L+1)%3], a dense level has
    // counted its winners.  An empty next level ends the search now.
    const bool dense_td = dir != DIR_BU && E > a.dense_edges;
    const unsigned long long next_cnt =
        dense_td ? a.ctl[CTL_DCNT0 + ((L + 1) & 1)] : a.ctl[CTL_RING0 + (L + 1) % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&a.ctl[CTL_MREACH], static_cast<unsigned long long>(E));
        a.ctl[CTL_EXP0 + ((L + 1) & 1)] = static_cast<unsigned long long>(explored);
        a.ctl[CTL_DCNT0 + (L & 1)] = 0;  // read by compact(L-1); counts level L+2 next
        if ((next_cnt >> 32) == 0) a.ctl[CTL_DONE] = 1;
    }
    if ((next_cnt >> 32) == 0) return;
    if (dir == DIR_BU) {
        // Back top-down when level L+1 is small and shrinking (a bottom-up
        // level costs a pass over every vertex), or gets no bottom-up launch.
        const unsigned long long nx = a.ctl[CTL_RING0 + (L + 1) % 3];
        const uint32_t F1 = static_cast<uint32_t>(nx >> 32);
        const bool to_top_down = !bu_next || go_top_down(a, F, F1);
        if (F1 == 0 || !to_top_down) return;  // empty, or level L+1 stays bottom-up
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl[CTL_REP0 + (L + 1) % 3], REP_QUEUE);
        for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
            compact_chunk(a, L, c, true);
            __syncthreads();
        }
        return;
    }
    if (E <= a.dense_edges) return;  // sparse level: expand() queued level L+1
    // Dense top-down level: expand(L) counted level L+1.  If level L+1 will
    // go bottom-up (the test choose_dir(L+1) will make), only its bitmap is
    // needed; otherwise the queue too.
    const unsigned long long nx = a.ctl[CTL_DCNT0 + ((L + 1) & 1)];
    const uint32_t F1 = static_cast<uint32_t>(nx >> 32), E1 = static_cast<uint32_t>(nx);
    if (bu_next && a.bottom_up && F1 > 0 && go_bottom_up(a, explored, E1, E)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.ctl[CTL_RING0 + (L + 1) % 3] = nx;
            a.ctl[CTL_REP0 + (L + 1) % 3] = REP_BITMAP;
        }
        uint32_t* fout = a.fb[(L + 1) % 3];
        for (uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x; i < a.nwords; i += gridDim.x * kCompactBlock) {
            const uint32_t vw = a.vis[i];
            fout[i] = vw & ~a.prev[i];
            a.prev[i] = vw;
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[CTL_REP0 + (L + 1) % 3] = REP_QUEUE | REP_BITMAP;
    for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
        compact_chunk(a, L, c, false);
        __syncthreads();
    }
}

// Deep, thin levels (road-like graphs: thousands of levels of a few thousand
// vertices).  A level's cost there is its chain of dependent memory round
// trips, not bytes, so these levels use a kernel with a short chain: a thread
// takes one frontier vertex (queue entries are bare vertex ids, loaded
// together with the level's counter), loads its row bounds, then its columns,
// claims them with atomicOr on the visited bitmap (no probe first: few
// claims per level), and each wave reserves its winners' queue slots with
// one atomicAdd -- five round trips against about nine of the edge-balanced
// kernel.  Rows longer than kDeepRow go to the wave, which claims 64 entries
// per step.  Level L's CSR entries are summed during its expansion
// (deepe[L & 1]) and booked by block 0 of level L+1; the search ends when a
// level finds its frontier empty.
constexpr int kDeepBlock = 256;
constexpr int kDeepRow = 8;

__global__ __launch_bounds__(kDeepBlock) void bfs_deep_kernel(BfsArgs a, int L) {
    const uint32_t lane = threadIdx.x & 63u;
    const int cur = L & 1, nxt = cur ^ 1;
    const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kDeepBlock + threadIdx.x;
    const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kDeepBlock;
    const uint64_t n = static_cast<uint64_t>(a.n);
    // This thread's first frontier entry is loaded beside the counters; it
    // is used only if it lies below the frontier size.
    const int32_t v_first = gtid < n ? a.q_v[cur][gtid] : -1;
    const unsigned long long stop = a.ctl[CTL_DONE] | a.ctl[CTL_ERR];
    const unsigned long long cnt = a.ctl[CTL_RING0 + L % 3];
    if (stop) return;
    const uint32_t F = static_cast<uint32_t>(cnt >> 32);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long e_prev = a.ctl[CTL_DEEPE0 + ((L + 1) & 1)];  // level L-1's entries
        a.ctl[CTL_DEEPE0 + ((L + 1) & 1)] = 0;                                 // level L+1 sums there next
        atomicAdd(&a.ctl[CTL_MREACH], e_prev);
        atomicMax(&a.ctl[CTL_MAXE], e_prev);
        if (F == 0) {
            a.ctl[CTL_DONE] = 1;
        } else {
            a.ctl[CTL_RING0 + (L + 2) % 3] = 0;  // held level L-1's counter; counts level L+2 next
            atomicAdd(&a.ctl[CTL_NREACH], static_cast<unsigned long long>(F));
            atomicMax(&a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L));
        }
    }
    if (F == 0) return;
    if (F > n) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
        return;
    }
    int32_t* const levels = a.p->levels;
    int32_t* const parents = a.p->parents;
    unsigned long long* const ring_next = &a.ctl[CTL_RING0 + (L + 1) % 3];
    bool bad = false;
    uint64_t my_edges = 0;
    // Reserves c queue slots for each lane (wave-uniform call, one atomicAdd
    // per wave); returns the lane's first slot, or ~0u if the queue would
    // overflow (reported as an error).
    auto reserve = [&](uint32_t c) -> uint32_t {
        const uint32_t incl = wave_inclusive_scan(c);
        const uint32_t tot = __shfl(incl, 63, 64);
        if (tot == 0) return 0u;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(ring_next, static_cast<unsigned long long>(tot) << 32);
        base = __shfl(base, 0, 64);
        if ((base >> 32) + tot > n) {
            bad = true;
            return ~0u;
        }
        return static_cast<uint32_t>(base >> 32) + incl - c;
    };
    for (uint64_t i = gtid;; i += gstride) {
        const bool has = i < F;
        if (!wave_ballot(has)) break;  // the lanes of a wave hold consecutive entries
        int32_t v = -1;
        if (has) {
            v = i == gtid ? v_first : a.q_v[cur][i];
            if (static_cast<uint64_t>(static_cast<uint32_t>(v)) >= n) {
                bad = true;
                v = -1;
            }
        }
        // Rows of at most 4 entries come whole from `head` (no offsets, no
        // column loads); longer rows, or graphs without `head`, from the CSR.
        int4 hd = make_int4(-1, -1, -1, -1);
        if (v >= 0 && a.head) hd = a.head[v];
        const bool from_csr = v >= 0 && (!a.head || hd.w < -1);
        int64_t rb = 0, re = 0;
        if (from_csr) {
            rb = a.offsets[v];
            re = a.offsets[v + 1];
        }
        const uint32_t deg = from_csr ? static_cast<uint32_t>(re - rb)
                                      : static_cast<uint32_t>((hd.x >= 0) + (hd.y >= 0) + (hd.z >= 0) + (hd.w >= 0));
        my_edges += deg;
        const bool short_row = v >= 0 && deg <= kDeepRow;
        int32_t u[kDeepRow];
#pragma unroll
        for (int j = 0; j < kDeepRow; ++j) {
            if (from_csr) u[j] = (short_row && static_cast<uint32_t>(j) < deg) ? a.cols[rb + j] : -1;
            else u[j] = j == 0 ? hd.x : j == 1 ? hd.y : j == 2 ? hd.z : j == 3 ? hd.w : -1;
            if (u[j] >= 0 && static_cast<uint64_t>(u[j]) >= n) {
                bad = true;
                u[j] = -1;
            }
        }
        uint32_t old[kDeepRow];
#pragma unroll
        for (int j = 0; j < kDeepRow; ++j)
            old[j] = u[j] >= 0 ? atomicOr(&a.vis[static_cast<uint32_t>(u[j]) >> 5], 1u << (u[j] & 31)) : ~0u;
        uint32_t won = 0;  // bit j: u[j] claimed by this lane
#pragma unroll
        for (int j = 0; j < kDeepRow; ++j) {
            if (u[j] >= 0 && !((old[j] >> (u[j] & 31)) & 1u)) {
                levels[u[j]] = L + 1;
                if (parents) parents[u[j]] = v;
                won |= 1u << j;
            }
        }
        const uint32_t slot0 = reserve(static_cast<uint32_t>(__popc(won)));
        if (won && slot0 != ~0u) {
            uint32_t r = 0;
#pragma unroll
            for (int j = 0; j < kDeepRow; ++j)
                if ((won >> j) & 1u) a.q_v[nxt][slot0 + r++] = u[j];
        }
        // Long rows, one at a time by the whole wave.
        uint64_t lm = wave_ballot(v >= 0 && !short_row);
        while (lm) {
            const int hl = __ffsll(static_cast<long long>(lm)) - 1;
            lm &= lm - 1;
            const int32_t hv = __shfl(v, hl, 64);
            const int64_t hrb = __shfl(rb, hl, 64), hre = __shfl(re, hl, 64);
            for (int64_t b = hrb; b < hre; b += 64) {
                int32_t uu = b + lane < hre ? a.cols[b + lane] : -1;
                if (uu >= 0 && static_cast<uint64_t>(uu) >= n) {
                    bad = true;
                    uu = -1;
                }
                const uint32_t o = uu >= 0 ? atomicOr(&a.vis[static_cast<uint32_t>(uu) >> 5], 1u << (uu & 31)) : ~0u;
                const bool w = uu >= 0 && !((o >> (uu & 31)) & 1u);
                if (w) {
                    levels[uu] = L + 1;
                    if (parents) parents[uu] = hv;
                }
                const uint32_t slot = reserve(w ? 1u : 0u);
                if (w && slot != ~0u) a.q_v[nxt][slot] = uu;
            }
        }
    }
    const uint64_t e = wave_sum(my_edges);
    if (lane == 0 && e) atomicAdd(&a.ctl[CTL_DEEPE0 + (L & 1)], static_cast<unsigned long long>(e));
    if (bad) atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
}

// The tail of a search (one launch per level, symmetric graphs).  Once the
// frontier is small, a launch per level costs more than the level's work
// (about 4.5 us of dispatch and kernel boundary), and since searches of one
// graph differ in depth, the levels queued for the deepest are no-ops for
// the others.  From the level the host has learned (the first level of the
// last run of small levels in recent searches), one workgroup runs the rest
// of the search with workgroup barriers only; no workgroup waits on another.
// A level's entries (vertex, first CSR entry, prefix of degrees) live in
// LDS.  Each thread takes kTailK edges per pass (owner by binary search in
// the prefix), claims their targets with atomicOr on the visited bitmap and
// queues the winners with their row bounds (one LDS reservation per wave),
// so the next level's size is known before it starts.
//  * A level that does not fit (more than kTailCap entries or kTailMaxEdges
//    CSR entries) is not expanded here.  At the first level nothing has
//    changed yet: the host goes on with the usual launches from that level
//    (GB_TAIL_DECLINED).  Later, the level is left as a bare vertex queue
//    and the host goes on with bfs_deep_kernel from there
//    (GB_TAIL_HANDED_BACK).
//  * The counters the host reads (reached vertices, CSR entries, deepest
//    level, largest frontier, done) are kept as expand() keeps them.
static_assert(kTailCap <= 2 * kTailBlock, "the prefix pass takes two entries per thread");

__device__ void tail_levels(const BfsArgs& a, int T) {
    __shared__ uint32_t s_pre[2][kTailCap + 1];  // exclusive prefix of degrees (+ the total)
    __shared__ int64_t s_rb[2][kTailCap];
    __shared__ int32_t s_v[2][kTailCap];
    __shared__ unsigned long long s_wave[kTailBlock / 64];
    __shared__ uint32_t s_count;
    __shared__ unsigned long long s_edges;
    __shared__ uint32_t s_bad;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if (a.ctl[CTL_DONE] | a.ctl[CTL_ERR]) return;
    const unsigned long long cnt = a.ctl[CTL_RING0 + T % 3];
    uint32_t F = static_cast<uint32_t>(cnt >> 32), E = static_cast<uint32_t>(cnt);
    if (F == 0) {
        if (tid == 0) {
            a.ctl[CTL_DONE] = 1;
            a.ctl[CTL_TAIL] = GB_TAIL_FINISHED | (static_cast<unsigned long long>(T) << 8);
        }
        return;
    }
    if (!(a.ctl[CTL_REP0 + T % 3] & REP_QUEUE) || F > static_cast<uint32_t>(kTailCap) || E > kTailMaxEdges) {
        if (tid == 0) a.ctl[CTL_TAIL] = GB_TAIL_DECLINED | (static_cast<unsigned long long>(T) << 8);
        return;
    }
    {
        const int q = T & 1;
        for (uint32_t i = tid; i < F; i += kTailBlock) {
            s_v[0][i] = a.q_v[q][i];
            s_rb[0][i] = a.q_rb[q][i];
            s_pre[0][i] = a.q_eoff[q][i];
        }
        if (tid == 0) {
            s_pre[0][F] = E;
            s_bad = 0;
            if (a.ctl[CTL_TAILFROM] == ~0ull) a.ctl[CTL_TAILFROM] = static_cast<unsigned long long>(T);
        }
    }
    int32_t* const levels = a.p->levels;
    int32_t* const parents = a.p->parents;
    const uint64_t n = static_cast<uint64_t>(a.n);
    for (int L = T;; ++L) {
        const int cur = (L - T) & 1, nxt = cur ^ 1;
        if (tid == 0) {
            a.ctl[CTL_NREACH] += F;
            a.ctl[CTL_MREACH] += E;
            a.ctl[CTL_MAXE] = std::max<unsigned long long>(a.ctl[CTL_MAXE], E);
            a.ctl[CTL_MAXLEV] = std::max<unsigned long long>(a.ctl[CTL_MAXLEV], static_cast<unsigned long long>(L));
            s_count = 0;
            s_edges = 0;
        }
        __syncthreads();
        bool bad = false;
        for (uint32_t e0 = 0; e0 < E; e0 += kTailBlock * kTailK) {
            int32_t u[kTailK];
            uint32_t own[kTailK];
#pragma unroll
            for (int j = 0; j < kTailK; ++j) {
                const uint32_t e = e0 + j * kTailBlock + tid;
                u[j] = -1;
                own[j] = 0;
                if (e < E) {
                    own[j] = owner_of(s_pre[cur], F, e);
                    const int64_t i = s_rb[cur][own[j]] + (e - s_pre[cur][own[j]]);
                    if (i < 0 || i >= a.m) {
                        bad = true;
                    } else {
                        const int32_t c = a.cols[i];
                        if (static_cast<uint64_t>(static_cast<uint32_t>(c)) >= n) bad = true;
                        else u[j] = c;
                    }
                }
            }
            uint32_t old[kTailK];
#pragma unroll
            for (int j = 0; j < kTailK; ++j)
                old[j] = u[j] >= 0 ? atomicOr(&a.vis[static_cast<uint32_t>(u[j]) >> 5], 1u << (u[j] & 31)) : ~0u;
            uint32_t won = 0;
#pragma unroll
            for (int j = 0; j < kTailK; ++j) {
                if (u[j] >= 0 && !((old[j] >> (u[j] & 31)) & 1u)) {
                    won |= 1u << j;
                    levels[u[j]] = L + 1;
                    if (parents) parents[u[j]] = s_v[cur][own[j]];
                }
            }
            // The winners' rows make the next level's entries.
            int64_t wrb[kTailK];
            uint32_t wdeg[kTailK];
            unsigned long long my_edges = 0;
#pragma unroll
            for (int j = 0; j < kTailK; ++j) {
                wrb[j] = 0;
                wdeg[j] = 0;
                if ((won >> j) & 1u) {
                    wrb[j] = a.offsets[u[j]];
                    wdeg[j] = static_cast<uint32_t>(a.offsets[u[j] + 1] - wrb[j]);
                    my_edges += wdeg[j];
                }
            }
            const uint32_t c = static_cast<uint32_t>(__popc(won));
            const uint32_t incl = wave_inclusive_scan(c);
            const uint32_t tot = __shfl(incl, 63, 64);
            if (tot == 0) continue;  // wave-uniform
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&s_count, tot);
            base = __shfl(base, 0, 64);
            const unsigned long long we = wave_sum(my_edges);
            if (lane == 0) atomicAdd(&s_edges, we);
            uint32_t slot = base + incl - c;
#pragma unroll
            for (int j = 0; j < kTailK; ++j) {
                if ((won >> j) & 1u) {
                    if (slot < static_cast<uint32_t>(kTailCap)) {
                        s_v[nxt][slot] = u[j];
                        s_rb[nxt][slot] = wrb[j];
                        s_pre[nxt][slot] = wdeg[j];  // the degree until the prefix pass
                    } else if (slot < n) {
                        a.q_v[(L + 1) & 1][slot] = u[j];  // beyond the LDS: kept for the deep kernel
                    } else {
                        bad = true;
                    }
                    ++slot;
                }
            }
        }
        if (bad) s_bad = 1;
        __syncthreads();
        const uint32_t Fn = s_count;
        const unsigned long long En = s_edges;
        if (s_bad) {
            if (tid == 0) atomicOr(&a.ctl[CTL_ERR], ERR_RANGE);
            return;
        }
        if (Fn == 0) {
            if (tid == 0) {
                a.ctl[CTL_DONE] = 1;
                a.ctl[CTL_TAIL] = GB_TAIL_FINISHED | (static_cast<unsigned long long>(L + 1) << 8);
            }
            return;
        }
        if (Fn > static_cast<uint32_t>(kTailCap) || En > kTailMaxEdges) {
            // Level L+1 goes on in bfs_deep_kernel, which reads a bare
            // vertex queue and counts the level itself.
            for (uint32_t i = tid; i < std::min<uint32_t>(Fn, kTailCap); i += kTailBlock)
                a.q_v[(L + 1) & 1][i] = s_v[nxt][i];
            if (tid == 0) {
                a.ctl[CTL_RING0 + (L + 1) % 3] = static_cast<unsigned long long>(Fn) << 32 |
                                                 std::min<unsigned long long>(En, 0xffffffffull);
                a.ctl[CTL_RING0 + (L + 2) % 3] = 0;
                a.ctl[CTL_DEEPE0] = 0;
                a.ctl[CTL_DEEPE0 + 1] = 0;
                a.ctl[CTL_TAILFROM] = ~0ull;
                a.ctl[CTL_TAIL] = GB_TAIL_HANDED_BACK | (static_cast<unsigned long long>(L + 1) << 8);
            }
            return;
        }
        // The next level's prefix of degrees, in place (two entries per thread).
        const uint32_t i0 = 2 * tid, i1 = 2 * tid + 1;
        const unsigned long long d0 = i0 < Fn ? s_pre[nxt][i0] : 0u;
        const unsigned long long d1 = i1 < Fn ? s_pre[nxt][i1] : 0u;
        unsigned long long total;
        const unsigned long long ex = block_exclusive_scan<kTailBlock>(d0 + d1, s_wave, total);
        if (i0 < Fn) s_pre[nxt][i0] = static_cast<uint32_t>(ex);
        if (i1 < Fn) s_pre[nxt][i1] = static_cast<uint32_t>(ex + d0);
        if (tid == 0) s_pre[nxt][Fn] = static_cast<uint32_t>(total);
        F = Fn;
        E = static_cast<uint32_t>(En);
        __syncthreads();
    }
}

// The control block as the host reads it after the search: copied into
// host-visible memory by the kernel that ends the launched sequence, so the
// sequence holds no copy node (a copy is a separate blit dispatch, and in a
// replayed graph it also delays the submission of the kernels behind it).
// Every thread of the workgroup calls this after its last update of ctl.
// The snapshot's CTL_T1 word is the device clock at the store: with CTL_T0
// it times the sequence on the device without events.  The host fills the
// snapshot with kSnapUnwritten before the launch and takes it as complete
// once no word holds that value any more, so the words need no ordering
// among themselves (no system-scope fence here; an aligned 8-byte store
// reaches host memory whole).
__device__ __forceinline__ unsigned long long snap_word(const unsigned long long* ctl, uint32_t i) {
    return i == CTL_T1 ? static_cast<unsigned long long>(wall_clock64())
                       : __hip_atomic_load(ctl + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void snapshot_ctl(unsigned long long* ctl, unsigned long long* out) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x < CTL_WORDS) out[threadIdx.x] = snap_word(ctl, threadIdx.x);
}

// `snap` (host-visible, may be null): where the final control block goes.
__global__ __launch_bounds__(kTailBlock) void bfs_tail_kernel(BfsArgs a, int T, unsigned long long* snap) {
    tail_levels(a, T);
    if (snap) snapshot_ctl(a.ctl, snap);
}

// The control block to host-visible memory when the sequence does not end
// with the tail kernel (earlier kernels' updates are complete at its start).
__global__ __launch_bounds__(64) void bfs_snapshot_kernel(const unsigned long long* ctl, unsigned long long* out) {
    if (threadIdx.x < CTL_WORDS) out[threadIdx.x] = snap_word(ctl, threadIdx.x);
}

// Upload of a symmetric graph: each row's neighbours ordered by descending
// degree bucket (file order inside a bucket).  A bottom-up level probes a row
// from its start and stops at the first frontier neighbour; the hubs, which
// are reached early and fill the big frontiers, now come first.  Levels do
// not depend on the order.  One wave per row, a stable counting sort over the
// buckets: one pass counts (LDS atomics), a wave scan turns the counts into
// bases, a second pass places 64 entries at a time, bucket by bucket.
constexpr int kRowSortBlock = 256;

__device__ __forceinline__ int hub_first_key(const int64_t* offsets, int32_t u) {
    const int64_t d = offsets[u + 1] - offsets[u];
    const int b = d <= 0 ? 0 : min(64 - __clzll(static_cast<long long>(d)), kDegreeBuckets - 1);
    return kDegreeBuckets - 1 - b;  // degree_bucket(d), highest first
}

__global__ __launch_bounds__(kRowSortBlock) void row_sort_kernel(const int64_t* __restrict__ offsets,
                                                                 const int32_t* __restrict__ cols,
                                                                 int32_t* __restrict__ out, int64_t n) {
    __shared__ uint32_t s_base[kRowSortBlock / 64][kDegreeBuckets];
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* base = s_base[threadIdx.x >> 6];
    const int64_t nwaves = int64_t(gridDim.x) * (blockDim.x / 64);
    for (int64_t v = int64_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); v < n; v += nwaves) {
        const int64_t rb = offsets[v], re = offsets[v + 1];
        if (re == rb) continue;
        if (lane < kDegreeBuckets) base[lane] = 0;
        wave_lds_sync();
        for (int64_t c0 = rb; c0 < re; c0 += 64) {
            const int64_t j = c0 + lane;
            if (j < re) atomicAdd(&base[hub_first_key(offsets, cols[j])], 1u);
        }
        wave_lds_sync();
        const uint32_t cnt = lane < kDegreeBuckets ? base[lane] : 0u;
        const uint32_t incl = wave_inclusive_scan(cnt);
        wave_lds_sync();
        if (lane < kDegreeBuckets) base[lane] = incl - cnt;
        wave_lds_sync();
        for (int64_t c0 = rb; c0 < re; c0 += 64) {
            const int64_t j = c0 + lane;
            const bool valid = j < re;
            const int32_t u = valid ? cols[j] : 0;
            const int key = valid ? hub_first_key(offsets, u) : -1;
            uint64_t todo = wave_ballot(valid);
            while (todo) {
                const int leader = __ffsll(static_cast<long long>(todo)) - 1;
                const int kk = __shfl(key, leader, 64);
                const uint64_t mk = wave_ballot(key == kk);
                const uint32_t b = base[kk];
                if (key == kk) out[rb + b + rank_below(mk)] = u;
                wave_lds_sync();
                if (static_cast<int>(lane) == leader) base[kk] = b + static_cast<uint32_t>(__popcll(mk));
                wave_lds_sync();
                todo &= ~mk;
            }
        }
    }
}

// The first 4 entries of every row side by side (-1 past the row's end); the
// fourth, u, is stored as -2 - u (below -1) when the row goes on, so a row of
// at most 4 entries is complete without its offsets.
__global__ void row_head_kernel(const int64_t* __restrict__ offsets, const int32_t* __restrict__ cols,
                                int4* __restrict__ head, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int64_t rb = offsets[v], d = offsets[v + 1] - rb;
        head[v] = make_int4(d > 0 ? cols[rb] : -1, d > 1 ? cols[rb + 1] : -1, d > 2 ? cols[rb + 2] : -1,
                            d > 4 ? -2 - cols[rb + 3] : d > 3 ? cols[rb + 3] : -1);
    }
}

// One launch before level 0 (a kernel boundary costs more than this work):
//  * levels = -1, parents = -1; visited and queued bitmaps = the isolated
//    vertices (symmetric graphs: nothing can reach them) or 0; frontier
//    bitmaps = 0 when bottom-up levels may run;
//  * level 0 = the source alone: claimed, level 0, parent = itself, the only
//    queue entry, and the control block reset.  The thread that clears an
//    entry of the source also makes the source's store to it.
__global__ __launch_bounds__(kClearBlock) void bfs_clear_kernel(BfsArgs a, BfsParams arg) {
    // The call's parameters: `arg` for a direct launch, else (a replayed
    // sequence, whose arguments are fixed) the host-visible block, one
    // scalar read per wave.  The later kernels read the device copy made
    // here.
    const unsigned long long t0 = static_cast<unsigned long long>(wall_clock64());
    const BfsParams prm = arg.levels ? arg : *a.hp;
    const int32_t s = prm.source;
    int32_t* const levels = prm.levels;
    int32_t* const parents = prm.parents;
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.p = prm;
    const int64_t sw = static_cast<uint32_t>(s) >> 5;
    const uint32_t sbit = 1u << (static_cast<uint32_t>(s) & 31u);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    // Level 0 in this kernel (below): the source's row, and in workgroup 0
    // the row bounds of its neighbours, are loaded before the clearing
    // stores, so that their latency overlaps them.
    const bool inline_l0 = (prm.flags & BFS_PARAM_INLINE_L0) != 0;
    const int64_t rb = a.offsets[s];
    const uint32_t d = static_cast<uint32_t>(a.offsets[s + 1] - rb);
    int32_t u = -1;
    if (inline_l0 && threadIdx.x < d && threadIdx.x < kInlineL0) u = a.cols[rb + threadIdx.x];
    int64_t urb = 0, ure = 0;
    if (inline_l0 && blockIdx.x == 0 && u >= 0 && u < a.n) {
        urb = a.offsets[u];
        ure = a.offsets[u + 1];
    }
    // Thread i clears entries 4i..4i+3 with 16-byte stores when the output
    // rows allow it (entry i otherwise), and bitmap word i.
    const bool vec = ((reinterpret_cast<uintptr_t>(levels) | reinterpret_cast<uintptr_t>(parents)) & 15u) == 0;
    const int64_t nitems = vec ? (a.n + 3) / 4 : a.n;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nitems; i += stride) {
        if (!vec) {
            levels[i] = i == s ? 0 : -1;
            if (parents) parents[i] = i == s ? s : -1;
        } else if (i * 4 + 3 < a.n) {
            int32_t lv[4], pv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                lv[k] = i * 4 + k == s ? 0 : -1;
                pv[k] = i * 4 + k == s ? s : -1;
            }
            *reinterpret_cast<int4*>(levels + i * 4) = make_int4(lv[0], lv[1], lv[2], lv[3]);
            if (parents) *reinterpret_cast<int4*>(parents + i * 4) = make_int4(pv[0], pv[1], pv[2], pv[3]);
        } else {
            for (int64_t v = i * 4; v < a.n; ++v) {
                levels[v] = v == s ? 0 : -1;
                if (parents) parents[v] = v == s ? s : -1;
            }
        }
        if (i < a.nwords) {
            const uint32_t w = (a.iso ? a.iso[i] : 0u) | (i == sw ? sbit : 0u);
            a.vis[i] = w;
            a.prev[i] = w;
            if (a.bottom_up) {
                a.fb[0][i] = 0u;
                a.fb[1][i] = 0u;
                a.fb[2][i] = 0u;
            }
        }
    }
    // Level 0 here, for a source of at most kInlineL0 neighbours on a
    // one-launch (symmetric) search, as expand(0) would run it: its launch
    // is a chain of dependent steps for one vertex.  Every workgroup reads
    // the source's row; once its own clearing stores are done, it stores
    // level 1 (and parent, visited and queued bits) for the neighbours whose
    // entries and bitmap words it owns; workgroup 0 builds the queue of
    // level 1 and the control block expand(0) would have left.
    __shared__ int32_t s_nb[kInlineL0];
    __shared__ unsigned long long s_wave[kClearBlock / 64];
    bool win = false;
    if (inline_l0) {
        if (threadIdx.x < kInlineL0) s_nb[threadIdx.x] = u;
        __threadfence_block();  // the clearing stores above come before the stores below
        __syncthreads();
        // a winner: a vertex, not the source, the first of its duplicates
        win = d <= kInlineL0 && u >= 0 && u < a.n && u != s;
        if (!(prm.flags & BFS_PARAM_ROWS_UNIQUE))
            for (uint32_t j = 0; win && j < threadIdx.x; ++j)
                if (s_nb[j] == u) win = false;
        if (win) {
            const int64_t li = vec ? (u >> 2) : u;  // the item whose thread cleared entry u
            if ((li % stride) / blockDim.x == blockIdx.x) {
                levels[u] = 1;
                if (parents) parents[u] = s;
            }
            const int64_t w = u >> 5;
            if ((w % stride) / blockDim.x == blockIdx.x) {
                const uint32_t bit = 1u << (u & 31);
                atomicOr(&a.vis[w], bit);
                atomicOr(&a.prev[w], bit);
                if (prm.flags & BFS_PARAM_BU_NEXT) atomicOr(&a.fb[1][w], bit);
            }
        }
    }
    if (blockIdx.x != 0) return;
    unsigned long long l1_total = 0;
    uint32_t l1_zero = 0;
    if (inline_l0) {
        const uint32_t ud = win ? static_cast<uint32_t>(ure - urb) : 0u;
        const bool keep = win && ud > 0;
        const unsigned long long ex =
            block_exclusive_scan<kClearBlock>(keep ? ((1ull << 32) | ud) : 0ull, s_wave, l1_total);
        if (keep) {
            const uint32_t slot = static_cast<uint32_t>(ex >> 32), eo = static_cast<uint32_t>(ex);
            a.q_rb[1][slot] = urb;
            a.q_eoff[1][slot] = eo;
            a.q_v[1][slot] = u;
            mark_tiles(a.tile_start[1], a.tile_edges, eo, ud, slot);
        }
        l1_zero = static_cast<uint32_t>(__syncthreads_count(win && ud == 0));  // reached, never queued
    }
    if (threadIdx.x == 0) {
        a.q_rb[0][0] = rb;
        a.q_eoff[0][0] = 0;
        a.q_v[0][0] = s;
        const unsigned long long err = inline_l0 && d > kInlineL0 ? ERR_RANGE : 0ull;  // never asked for
        for (int i = 0; i < CTL_WORDS; ++i) a.ctl[i] = 0;
        a.ctl[CTL_ERR] = err;
        a.ctl[CTL_RING0] = d ? ((1ull << 32) | d) : 0ull;
        a.ctl[CTL_NREACH] = d ? 0 : 1;  // expand(0) counts a source with edges
        a.ctl[CTL_REP0] = REP_QUEUE;
        a.ctl[CTL_BUFIRST] = ~0ull;
        a.ctl[CTL_TAILFROM] = ~0ull;
        a.ctl[CTL_T0] = t0;
        if (inline_l0) {
            // what expand(0) leaves: level 0 (F = 1, E = d) counted and
            // top-down, level 1 queued
            a.ctl[CTL_NREACH] = 1 + l1_zero;
            a.ctl[CTL_MAXLEV] = l1_zero ? 1 : 0;
            a.ctl[CTL_MAXE] = d;
            a.ctl[CTL_DIR0] = DIR_TD;
            a.ctl[CTL_REP0 + 1] = (prm.flags & BFS_PARAM_BU_NEXT) ? (REP_QUEUE | REP_BITMAP) : REP_QUEUE;
            a.ctl[CTL_MREACH] = d;
            a.ctl[CTL_EXP0 + 1] = d;
            a.ctl[CTL_FCNT0] = a.ctl[CTL_RING0];
            a.ctl[CTL_RCNT0] = 1;
            a.ctl[CTL_TAILFROM] = (d <= kTailLearnEdges && 1 <= kTailCap) ? 0ull : ~0ull;
            a.ctl[CTL_RING0 + 1] = l1_total;
        }
    }
    for (uint64_t tt = threadIdx.x; tt * a.tile_edges < d; tt += blockDim.x) a.tile_start[0][tt] = 0;
}

using ExpandFn = void (*)(BfsArgs, int, int, int, int);

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

int min_tile_edges() {
    int t = kKernels[0].block * kKernels[0].k;
    for (const auto& ks : kKernels) t = std::min(t, ks.block * ks.k);
    return t;
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
    symmetric = is_symmetric(g);
    rows_unique = !(g.flags & GB_KEEP_ORDER);
    deg8.resize(static_cast<size_t>(n));
    for (int64_t v = 0; v < n; ++v) {
        const int64_t dv = g.offsets[v + 1] - g.offsets[v];
        n_edged += dv > 0 ? 1 : 0;
        deg8[v] = static_cast<uint8_t>(std::min<int64_t>(dv, 255));
    }
    try {
        GG_HIP(hipMalloc(&offsets, sizeof(int64_t) * (n + 1)));
        GG_HIP(hipMalloc(&cols, sizeof(int32_t) * std::max<int64_t>(m, 1)));
        graph_bytes = sizeof(int64_t) * (n + 1) + sizeof(int32_t) * std::max<int64_t>(m, 1);
        GG_HIP(hipMemcpy(offsets, g.offsets.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
        if (m) GG_HIP(hipMemcpy(cols, g.cols.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
        if (symmetric) {
            // Degree-0 vertices of a symmetric graph have no in-edges either:
            // the searches mark them visited up front.
            const int64_t nwords = (n + 31) / 32;
            std::vector<uint32_t> h(nwords, 0u);
            for (int64_t v = 0; v < n; ++v)
                if (g.offsets[v + 1] == g.offsets[v]) h[v >> 5] |= 1u << (v & 31);
            GG_HIP(hipMalloc(&iso, sizeof(uint32_t) * nwords));
            graph_bytes += sizeof(uint32_t) * nwords;
            GG_HIP(hipMemcpy(iso, h.data(), sizeof(uint32_t) * nwords, hipMemcpyHostToDevice));
            // Rows hub-first for the bottom-up levels (the host graph keeps
            // its order).
            const uint32_t rgrid = static_cast<uint32_t>(std::min<int64_t>((n + 3) / 4, int64_t(num_cus) * 8));
            if (m) {
                int32_t* sorted = nullptr;
                GG_HIP(hipMalloc(&sorted, sizeof(int32_t) * m));
                hipLaunchKernelGGL(row_sort_kernel, dim3(rgrid), dim3(kRowSortBlock), 0, nullptr, offsets, cols,
                                   sorted, n);
                const hipError_t e1 = hipGetLastError();
                const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(nullptr) : e1;
                if (e2 != hipSuccess) {
                    (void)hipFree(sorted);
                    GG_HIP(e2);
                }
                std::swap(cols, sorted);
                GG_HIP(hipFree(sorted));
            }
            // The first neighbours of every vertex side by side: a bottom-up
            // level reads them for 64 consecutive vertices in one coalesced
            // load instead of 64 scattered row starts.
            GG_HIP(hipMalloc(&head, sizeof(int4) * n));
            graph_bytes += sizeof(int4) * n;
            hipLaunchKernelGGL(row_head_kernel, dim3(static_cast<uint32_t>(std::min<int64_t>((n + 255) / 256,
                                                                                          int64_t(num_cus) * 8))),
                               dim3(256), 0, nullptr, offsets, cols, head, n);
            GG_HIP(hipGetLastError());
            GG_HIP(hipStreamSynchronize(nullptr));
        }
        build_ms_graph(*this, g);
        runner = std::make_unique<Runner>(*this);
    } catch (...) {
        release();  // members are then destroyed by the unwinding, once
        throw;
    }
}

DeviceGraph::~DeviceGraph() { release(); }

void DeviceGraph::release() {
    (void)hipSetDevice(device);
    runner.reset();
    if (offsets) (void)hipFree(offsets);
    if (cols) (void)hipFree(cols);
    if (iso) (void)hipFree(iso);
    if (head) (void)hipFree(head);
    for (void* p : {static_cast<void*>(ms_offsets), static_cast<void*>(ms_cols), static_cast<void*>(ms_perm),
                    static_cast<void*>(ms_iperm), static_cast<void*>(ms_heavy_tiles)})
        if (p) (void)hipFree(p);
    ms_offsets = nullptr;
    ms_cols = nullptr;
    ms_perm = ms_iperm = nullptr;
    ms_heavy_tiles = nullptr;
    offsets = nullptr;
    cols = nullptr;
    iso = nullptr;
    head = nullptr;
}

Runner::Runner(DeviceGraph& graph) : Runner(graph, nullptr) {}

Runner::Runner(DeviceGraph& graph, hipStream_t lane_stream)
    : g(graph), device(graph.device), n(graph.n), m(graph.m) {
    DeviceGuard guard(device);
    try {
        if (lane_stream) {
            stream = lane_stream;
            own_stream = false;
        } else {
            GG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        }
        GG_HIP(hipEventCreate(&ev_start));
        GG_HIP(hipEventCreate(&ev_stop));
        GG_HIP(hipEventCreate(&ev_kernel));
        GG_HIP(hipEventCreateWithFlags(&ev_batch[0], hipEventDisableTiming));
        GG_HIP(hipEventCreateWithFlags(&ev_batch[1], hipEventDisableTiming));
        // Coherent host memory: the kernels that end a launched sequence
        // store the control block here, and the host reads it after the
        // sequence's event.
        GG_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_ctl), sizeof(unsigned long long) * 2 * kCtlWords,
                             hipHostMallocCoherent | hipHostMallocMapped));
        GG_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&h_ctl_dev), h_ctl, 0));
        int khz = 0;
        GG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
        if (khz <= 0) fail(GB_EDEVICE, "device reports no wall clock rate");
        wall_khz = khz;
    } catch (...) {
        release();
        throw;
    }
}

Runner::~Runner() { release(); }

void Runner::release() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    free_workspace();
    free_ms_workspace();
    if (h_ctl) (void)hipHostFree(h_ctl);
    h_ctl = nullptr;
    h_ctl_dev = nullptr;
    for (hipEvent_t* e : {&ev_start, &ev_stop, &ev_kernel, &ev_batch[0], &ev_batch[1]}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    if (stream && own_stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

void Runner::free_workspace() {
    if (ws.graph_exec) (void)hipGraphExecDestroy(ws.graph_exec);
    void* ptrs[] = {ws.vis,    ws.prev,    ws.fb[0],  ws.fb[1],  ws.fb[2],  ws.q_rb[0],       ws.q_rb[1],
                    ws.q_eoff[0], ws.q_eoff[1], ws.q_v[0], ws.q_v[1], ws.ctl, ws.levels, ws.parents,
                    ws.tile_start[0], ws.tile_start[1], ws.params};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (ws.h_params) (void)hipHostFree(ws.h_params);
    ws = BfsWorkspace{};
}

void Runner::ensure_workspace() {
    // The tile indices are sized for the smallest tile of any supported
    // kernel shape, so a change of shape (the autotuner makes many) never
    // reallocates a buffer; only the captured sequence is rebuilt.
    const int T = min_tile_edges();
    const int64_t tile_cap = (m + T - 1) / T + 2;
    if (ws.ctl) return;
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
        alloc(&ws.fb[0], sizeof(uint32_t) * nwords);
        alloc(&ws.fb[1], sizeof(uint32_t) * nwords);
        alloc(&ws.fb[2], sizeof(uint32_t) * nwords);
        for (int i = 0; i < 2; ++i) {
            alloc(&ws.q_rb[i], sizeof(int64_t) * n);
            alloc(&ws.q_eoff[i], sizeof(uint32_t) * n);
            alloc(&ws.q_v[i], sizeof(int32_t) * n);
            alloc(&ws.tile_start[i], sizeof(uint32_t) * tile_cap);
        }
        alloc(&ws.ctl, sizeof(unsigned long long) * kCtlWords);
        alloc(&ws.levels, sizeof(int32_t) * n);
        alloc(&ws.parents, sizeof(int32_t) * n);
        alloc(&ws.params, sizeof(BfsParams));
        // Coherent host memory, read by the clear kernel of every search.
        GG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ws.h_params), sizeof(BfsParams),
                             hipHostMallocCoherent | hipHostMallocMapped));
        GG_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&ws.h_params_dev), ws.h_params, 0));
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
    a.fb[0] = ws.fb[0];
    a.fb[1] = ws.fb[1];
    a.fb[2] = ws.fb[2];
    a.iso = g.iso;
    a.head = g.head;
    for (int i = 0; i < 2; ++i) {
        a.q_rb[i] = ws.q_rb[i];
        a.q_eoff[i] = ws.q_eoff[i];
        a.q_v[i] = ws.q_v[i];
        a.tile_start[i] = ws.tile_start[i];
    }
    a.ctl = ws.ctl;
    a.p = ws.params;
    a.hp = ws.h_params_dev;
    a.n = n;
    a.m = m;
    a.nwords = static_cast<uint32_t>((n + 31) / 32);
    a.nchunks = static_cast<uint32_t>(ws.nchunks);
    a.dense_edges = static_cast<uint32_t>(std::min<int64_t>(g.cfg.dense_edges, 0xffffffffll));
    a.grow_until = bu_learned ? bu_lo + 1 : 0;  // the first learned bottom-up level (the window has a margin of one)
    a.tile_edges = static_cast<uint32_t>(g.cfg.block * g.cfg.edges_per_thread);
    a.bottom_up = (g.symmetric && g.cfg.direction_opt) ? 1u : 0u;
    a.alpha = static_cast<uint32_t>(g.cfg.alpha);
    a.beta = static_cast<uint32_t>(g.cfg.beta);

    if (occ_kernel != reinterpret_cast<const void*>(ks->expand)) {  // queried once per kernel shape
        int o = 0;
        GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, reinterpret_cast<const void*>(ks->expand), g.cfg.block, 0));
        if (o < 1) fail(GB_EDEVICE, "BFS kernel cannot be resident on a CU");
        occ_kernel = reinterpret_cast<const void*>(ks->expand);
        occ_blocks = o;
    }
    const int occ = occ_blocks;
    const int waves_per_block = (g.cfg.block + 63) / 64;
    const int default_per_cu = std::max(1, std::min(occ, 16 / waves_per_block));
    const int per_cu = g.cfg.blocks_per_cu > 0 ? std::min(g.cfg.blocks_per_cu, occ) : default_per_cu;
    const uint32_t grid = static_cast<uint32_t>(per_cu * g.num_cus);
    // compact(): chunks are independent; up to 2 workgroups of 1024 per CU.
    const uint32_t cgrid = static_cast<uint32_t>(std::min<int64_t>(ws.nchunks, int64_t(g.num_cus) * 2));
    const uint32_t clear_grid = static_cast<uint32_t>(std::min<int64_t>((n + 255) / 256, int64_t(g.num_cus) * 8));
    const uint32_t deep_grid = static_cast<uint32_t>(
        std::max<int64_t>(1, std::min<int64_t>((n + kDeepBlock - 1) / kDeepBlock, int64_t(g.num_cus) * 2)));
    // Levels [bu_lo, bu_hi], learned from earlier searches of this graph, may
    // go bottom-up; other levels run top-down.
    auto bu_at = [&](int L) { return a.bottom_up && L >= bu_lo && L <= bu_hi; };
    // Deep mode: graphs whose last searches were deep with thin frontiers
    // only (no bottom-up window left, no dense level) run one launch per
    // level, all levels sparse top-down.
    const bool deep = g.cfg.deep == GB_DEEP_ON ||
                      (g.cfg.deep == GB_DEEP_AUTO && bu_hi < bu_lo && depth_hint >= kDeepDepth);
    last_deep = deep;
    // Symmetric graphs outside deep mode run one launch per level: the
    // expansion keeps compact()'s counters, and the frontier forms the next
    // level needs are built by the level itself.
    a.fused = (a.bottom_up && !deep) ? 1u : 0u;
    a.n_edged = g.n_edged;
    last_fused = a.fused != 0;
    bool force_deep = false;  // the levels after a tail that handed back a large level
    auto launch_level = [&](int L) {
        const int bu = bu_at(L) ? 1 : 0, bu1 = bu_at(L + 1) ? 1 : 0, bu2 = bu_at(L + 2) ? 1 : 0;
        if (force_deep) {
            hipLaunchKernelGGL(bfs_deep_kernel, dim3(deep_grid), dim3(kDeepBlock), 0, stream, a, L);
            return;
        }
        // One launch per level in deep mode, and for level 0 (the source
        // alone, top-down; the clears have zeroed both frontier bitmaps):
        // expand(L) builds level L+1's queue and keeps compact(L)'s counters.
        if (a.fused) {
            hipLaunchKernelGGL(ks->expand, dim3(grid), dim3(g.cfg.block), 0, stream, a, L, bu, bu1, 0);
            return;
        }
        if (deep) {
            hipLaunchKernelGGL(bfs_deep_kernel, dim3(deep_grid), dim3(kDeepBlock), 0, stream, a, L);
            return;
        }
        if (L == 0) {
            hipLaunchKernelGGL(ks->expand, dim3(grid), dim3(g.cfg.block), 0, stream, a, L, 0, bu1, 1);
            return;
        }
        hipLaunchKernelGGL(ks->expand, dim3(grid), dim3(g.cfg.block), 0, stream, a, L, bu, bu1, 0);
        hipLaunchKernelGGL(bfs_compact_kernel, dim3(cgrid), dim3(kCompactBlock), 0, stream, a, L, bu1, bu2);
    };
    // The call's parameters: kernel arguments of a direct launch of the
    // clear kernel, or the host-visible block it reads when a replayed graph
    // reaches it (the previous call has ended, so nothing reads the block
    // meanwhile).
    ws.h_params->source = source;
    ws.h_params->flags = 0;
    ws.h_params->levels = d_levels;
    ws.h_params->parents = d_parents;
    // A captured launch takes its parameters from ws.h_params when it runs.
    auto launch_prologue = [&](bool captured, int32_t flags = 0) {
        BfsParams arg = *ws.h_params;
        if (captured) arg.levels = nullptr;
        arg.flags = flags;
        hipLaunchKernelGGL(bfs_clear_kernel, dim3(clear_grid), dim3(kClearBlock), 0, stream, a, arg);
    };

    int L = 0;
    int last_slot = 0;  // h_ctl slot holding the final control block
    // The first batch of levels ends with the tail kernel at level tail_at
    // (>= 0), which runs the rest of the search in one workgroup.  The
    // launches end with the control block stored to h_ctl slot `slot`: by the
    // tail kernel, or by the snapshot kernel.
    auto launch_levels = [&](int count, int tail_at, int slot) {
        unsigned long long* snap = h_ctl_dev + slot * kCtlWords;
        for (int i = 0; i < count; ++i) {
            if (L == tail_at) {
                hipLaunchKernelGGL(bfs_tail_kernel, dim3(1), dim3(kTailBlock), 0, stream, a, L, snap);
                ++L;
                return;
            }
            launch_level(L++);
        }
        hipLaunchKernelGGL(bfs_snapshot_kernel, dim3(1), dim3(64), 0, stream,
                           static_cast<const unsigned long long*>(ws.ctl), snap);
    };
    bool clock_timed = false;  // timed by the device clock stamps, not by ev_start / ev_stop
    // The host waits for a batch of levels by polling its snapshot: filled
    // with kSnapUnwritten here before the batch is enqueued, complete once
    // the kernel has overwritten every word (an event would add its record
    // and signal latency, a few microseconds each).  The stream is queried
    // now and then, so that a failed launch ends the wait with its error.
    auto arm = [&](int slot) {
        for (int i = 0; i < kCtlWords; ++i) __atomic_store_n(&h_ctl[slot * kCtlWords + i], kSnapUnwritten, __ATOMIC_RELAXED);
        std::atomic_thread_fence(std::memory_order_release);
    };
    auto written = [&](int slot) {
        const unsigned long long* w = &h_ctl[slot * kCtlWords];
        if (__atomic_load_n(w + CTL_T1, __ATOMIC_ACQUIRE) == kSnapUnwritten) return false;
        for (int i = 0; i < kCtlWords; ++i)
            if (__atomic_load_n(w + i, __ATOMIC_ACQUIRE) == kSnapUnwritten) return false;
        return true;
    };
    auto wait_snapshot = [&](int slot) {
        for (uint32_t it = 1; !written(slot); ++it) {
            if (it % 65536 != 0) continue;  // every 4,096 polls cost deep searches 0.6%
            const hipError_t r = hipStreamQuery(stream);
            if (r == hipErrorNotReady) continue;
            GG_HIP(r);
            if (!written(slot)) fail(GB_EDEVICE, "internal error: the launched levels stored no control block");
        }
    };
    auto enqueue = [&](int count, int slot, int tail_at = -1) {
        arm(slot);
        launch_levels(count, tail_at, slot);
        GG_HIP(hipGetLastError());
        if (!clock_timed) GG_HIP(hipEventRecord(ev_stop, stream));  // for the timing only
        last_slot = slot;
    };
    auto done = [&](int slot) {
        wait_snapshot(slot);
        return h_ctl[slot * kCtlWords + CTL_DONE] != 0 || h_ctl[slot * kCtlWords + CTL_ERR] != 0;
    };
    // Deeper than the first batch: batches of doubling size (at most 64
    // levels) are pipelined -- the done flag of batch b-1 is checked while
    // batch b runs.  Extra levels are no-ops on an empty frontier.
    auto continue_levels = [&]() {
        if (done(last_slot)) return;
        int batch = 8;
        int slot = last_slot ^ 1;
        enqueue(batch, slot);
        for (;;) {
            if (L > n + 64) fail(GB_EDEVICE, "internal error: more BFS levels than vertices");
            batch = std::min(batch * 2, 64);
            slot ^= 1;
            enqueue(batch, slot);
            if (done(slot ^ 1)) break;
        }
    };

    bool timed_traversal = true;  // ev_kernel marks the end of the clears
    last_tail_at = -1;
    if (g.cfg.mode == GB_MODE_LEVEL_LOOP) {
        // One level at a time; the host reads the next frontier size.
        GG_HIP(hipEventRecord(ev_start, stream));
        launch_prologue(false);
        GG_HIP(hipGetLastError());
        GG_HIP(hipEventRecord(ev_kernel, stream));
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
        GG_HIP(hipEventRecord(ev_stop, stream));
        GG_HIP(hipMemcpyAsync(h_ctl, ws.ctl, sizeof(unsigned long long) * kCtlWords, hipMemcpyDeviceToHost,
                              stream));
        GG_HIP(hipEventRecord(ev_batch[0], stream));
        GG_HIP(hipEventSynchronize(ev_batch[0]));
    } else {
        // Levels are queued ahead with no host synchronisation.  The first
        // batch covers the depth of recent searches + 1 (one empty level ends
        // the search).  Each batch ends with a snapshot of the control block;
        // the last batch's snapshot is the final state, so the common case
        // (the first batch reaches the end) needs one synchronisation.
        int first = std::max(depth_hint, 2);
        // One workgroup finishes the search from the tail level (learned, or
        // set), when that comes before the end of the first batch.
        const int tail_want = g.cfg.tail >= 0 ? g.cfg.tail : g.cfg.tail == GB_TAIL_AUTO ? tail_hint : -1;
        if (a.fused && g.cfg.tail >= 0) first = std::max(first, g.cfg.tail + 1);
        // At most kMaxCapturedLevels levels before the first done check, in
        // either mode: deeper searches go on in pipelined batches, so a
        // search shallower than the depth hint of a deep graph (thousands of
        // levels) runs at most one batch of no-op levels past its end.
        const int D = std::min(first, kMaxCapturedLevels);
        const int tail_at = a.fused && tail_want >= 0 && tail_want < D ? tail_want : -1;
        last_tail_at = tail_at;
        if (g.cfg.mode == GB_MODE_GRAPH) {
            // The parameter upload, the clears and the first levels are one
            // captured hipGraph, replayed while the depth, the bottom-up
            // window and the kernel shape stay the same.
            const bool reuse = ws.graph_exec && ws.g_depth == D && ws.g_lo == bu_lo && ws.g_hi == bu_hi &&
                               ws.g_block == g.cfg.block && ws.g_k == g.cfg.edges_per_thread &&
                               ws.g_dense == int64_t(a.dense_edges) && ws.g_bu == int(a.bottom_up) &&
                               ws.g_alpha == int(a.alpha) && ws.g_beta == int(a.beta) && ws.g_deep == int(deep) &&
                               ws.g_tail == tail_at;
            if (!reuse) {
                if (ws.graph_exec) (void)hipGraphExecDestroy(ws.graph_exec);
                ws.graph_exec = nullptr;
                hipGraph_t captured = nullptr;
                GG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                try {
                    launch_prologue(true);
                    launch_levels(D, tail_at, 0);
                    GG_HIP(hipGetLastError());
                } catch (...) {
                    hipGraph_t dummy = nullptr;
                    (void)hipStreamEndCapture(stream, &dummy);
                    if (dummy) (void)hipGraphDestroy(dummy);
                    throw;
                }
                GG_HIP(hipStreamEndCapture(stream, &captured));
                hipError_t e = hipGraphInstantiate(&ws.graph_exec, captured, nullptr, nullptr, 0);
                (void)hipGraphDestroy(captured);
                GG_HIP(e);
                ws.g_depth = D;
                ws.g_lo = bu_lo;
                ws.g_hi = bu_hi;
                ws.g_block = g.cfg.block;
                ws.g_k = g.cfg.edges_per_thread;
                ws.g_dense = int64_t(a.dense_edges);
                ws.g_bu = int(a.bottom_up);
                ws.g_alpha = int(a.alpha);
                ws.g_beta = int(a.beta);
                ws.g_deep = int(deep);
                ws.g_tail = tail_at;
            }
            arm(0);
            GG_HIP(hipEventRecord(ev_start, stream));
            GG_HIP(hipGraphLaunch(ws.graph_exec, stream));
            GG_HIP(hipEventRecord(ev_stop, stream));
            L = tail_at >= 0 ? tail_at + 1 : D;  // the level after the captured launches
            last_slot = 0;
        } else {
            // Direct launches (the default): for a sequence of about ten
            // kernels a replayed graph costs more device and host time here
            // than the launches themselves.  No start event: recording one
            // ahead of the clear kernel kept the GPU idle for its host time
            // (about 4 us per search); the device time comes from the clock
            // stamps of the first and the last kernel instead.
            clock_timed = true;
            // A source of at most kInlineL0 neighbours: the clear kernel runs
            // level 0 (on one-launch searches, where it is top-down, not
            // dense, and not the tail's), and the launches start at level 1.
            const uint32_t ds = g.deg8[static_cast<size_t>(source)];
            const uint32_t dense_min0 = (1 <= a.grow_until) ? kGrowDenseEdges : a.dense_edges;
            const bool inline_l0 = a.fused && tail_at != 0 && ds >= 1 && ds <= kInlineL0 &&
                                   !(ds > dense_min0 && bu_at(1));
            launch_prologue(false, inline_l0 ? (BFS_PARAM_INLINE_L0 | (bu_at(1) ? BFS_PARAM_BU_NEXT : 0) |
                                                (g.rows_unique ? BFS_PARAM_ROWS_UNIQUE : 0))
                                             : 0);
            if (inline_l0) L = 1;
            enqueue(inline_l0 ? D - 1 : D, 0, tail_at);
        }
        timed_traversal = false;
        if (tail_at >= 0 && !done(0)) {
            // The tail declined its first level (state untouched: the usual
            // launches go on from there) or handed back a large level as a
            // bare queue (the deep kernel goes on from there).
            const unsigned long long tw = h_ctl[CTL_TAIL];
            if ((tw & 0xffu) == GB_TAIL_HANDED_BACK) {
                L = static_cast<int>(tw >> 8);
                force_deep = true;
            } else {
                L = tail_at;
            }
        }
        continue_levels();
        wait_snapshot(last_slot);
        if (!clock_timed) GG_HIP(hipEventSynchronize(ev_stop));
    }
    const unsigned long long* fin = h_ctl + last_slot * kCtlWords;
    if (fin[CTL_ERR] & ERR_TILE) fail(GB_EDEVICE, "internal error: inconsistent frontier tile index");
    if (fin[CTL_ERR] & ERR_RANGE) fail(GB_EDEVICE, "internal error: frontier entry or column index out of range");
    if (!fin[CTL_DONE]) fail(GB_EDEVICE, "internal error: BFS did not reach an empty frontier");
    last_num_levels = static_cast<int>(fin[CTL_MAXLEV]) + 1;
    last_tail_outcome = last_tail_at >= 0 ? static_cast<int>(fin[CTL_TAIL] & 0xffu) : 0;
    last_bu_first = fin[CTL_BULAST] >= fin[CTL_BUFIRST] && fin[CTL_BUFIRST] != ~0ull ? int(fin[CTL_BUFIRST]) : -1;
    last_bu_last = last_bu_first >= 0 ? int(fin[CTL_BULAST]) : -1;
    learn_from_search(static_cast<int64_t>(fin[CTL_MAXE]) > static_cast<int64_t>(a.dense_edges), fin[CTL_TAILFROM]);
    float ms = 0.f;
    if (clock_timed) {
        // The clear kernel's start to the final snapshot, at the device
        // clock's rate (kHz: ticks per ms).
        const int64_t ticks = static_cast<int64_t>(fin[CTL_T1] - fin[CTL_T0]);
        ms = ticks > 0 ? static_cast<float>(static_cast<double>(ticks) / wall_khz) : 0.f;
    } else {
        GG_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
    }
    last_ms = ms;
    if (st) {
        float kms = ms;
        if (timed_traversal) GG_HIP(hipEventElapsedTime(&kms, ev_kernel, ev_stop));
        st->ms = ms;
        st->kernel_ms = kms;
        st->num_levels = last_num_levels;
        st->reserved = 0;
        st->n_reached = static_cast<int64_t>(fin[CTL_NREACH]);
        st->m_reached = static_cast<int64_t>(fin[CTL_MREACH]);
        st->copy_ms = 0;
    }
}

// Launch hints for the next search of this graph, from the one just done:
//  * depth_hint: levels queued before the first done check -- the deepest
//    recent search + 1, lowered after 256 shallower searches in a row (a
//    change re-captures the replayed sequence, which costs far more host
//    time than the no-op launches of one spare level);
//  * bottom-up window: grown to cover the levels this search went bottom-up
//    (the first learned window replaces the default); after two searches in
//    a row with only small frontiers (deep, thin graphs) no bottom-up
//    launches at all, until a search has a dense level again.
//  * tail_hint (one launch per level only): the level from which one
//    workgroup runs the rest of the search -- the level by which 95% of the
//    last kTailWindow searches had started their final run of small levels
//    (the latest start, while fewer than 20 searches are recorded).  A
//    search whose run starts later has its tail declined at that level and
//    goes on with the usual launches after a host round trip; the others
//    each save a launch.  On the R-MAT s22 sources the runs start at level
//    5 (21 of 64), 6 (42) and 7 (1).
// All only shape launches; any level without a bottom-up launch runs
// top-down, and a tail that meets a large level hands it back.
void Runner::learn_from_search(bool had_dense_level, unsigned long long tail_from) {
    if (last_fused) {
        const int t = tail_from >= static_cast<unsigned long long>(kNoTail) ? kNoTail : static_cast<int>(tail_from);
        tail_seen[tail_seen_n % kTailWindow] = t;
        ++tail_seen_n;
        const int cnt = std::min(tail_seen_n, kTailWindow);
        int sorted[kTailWindow];
        std::copy(tail_seen, tail_seen + cnt, sorted);
        std::sort(sorted, sorted + cnt);
        tail_hint = sorted[(cnt * 95 + 99) / 100 - 1];  // the ceil(0.95 cnt)-th smallest
    }
    // compact() of the last level sets the done flag; with one launch per
    // level, expand() of the empty level after it does.
    const int want = last_num_levels + (last_fused ? 1 : 0);
    if (want >= depth_hint) {
        depth_hint = want;
        short_streak = 0;
    } else if (++short_streak >= 256) {
        depth_hint = want;
        short_streak = 0;
    }
    if (last_bu_first >= 0) {
        thin_streak = 0;
        if (bu_learned && bu_hi >= bu_lo) {
            bu_lo = std::min(bu_lo, last_bu_first);
            bu_hi = std::max(bu_hi, last_bu_last + 2);
        } else {
            bu_lo = std::max(1, last_bu_first - 1);
            bu_hi = last_bu_last + 2;
        }
        bu_lo = std::max(1, std::min(bu_lo, last_bu_first - 1));  // one level of margin below
        bu_learned = true;
    } else if (had_dense_level) {
        thin_streak = 0;
        if (bu_hi < bu_lo) {  // re-arm the default window
            bu_lo = 1;
            bu_hi = kDefaultBuHigh;
            bu_learned = false;
        }
    } else if (++thin_streak >= 2) {
        bu_lo = 1;
        bu_hi = 0;
    }
}

void Runner::bfs_to_host(int32_t source, int32_t* h_levels, int32_t* h_parents, gb_stats* st) {
    if (!h_levels) fail(GB_EINVAL, "levels buffer is null");
    DeviceGuard guard(device);
    ensure_workspace();
    bfs(source, ws.levels, h_parents ? ws.parents : nullptr, st);
    // The search has ended (bfs() waited for it): the events time the copy alone.
    GG_HIP(hipEventRecord(ev_start, stream));
    GG_HIP(hipMemcpyAsync(h_levels, ws.levels, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
    if (h_parents) GG_HIP(hipMemcpyAsync(h_parents, ws.parents, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
    GG_HIP(hipEventRecord(ev_stop, stream));
    GG_HIP(hipStreamSynchronize(stream));
    if (st) {
        float cms = 0.f;
        GG_HIP(hipEventElapsedTime(&cms, ev_start, ev_stop));
        st->copy_ms = cms;
    }
}

}  // namespace gg
