// Batched multi-source BFS: up to 64 sources per launch sequence, one bit
// per source in 64-bit masks (MS-BFS, Then et al., VLDB 2015).  There is no
// counterpart in the reference, whose engine runs one source per process
// (1D Partition/src/gpuBFS.cu:13-116); each source's levels must equal what
// the single-source path (and cpuBFS, 1D Partition/src/cpuBFS.cpp:5-62)
// gives for that source.
//
// Per vertex: seen (sources that reached it), next (sources reaching it in
// the level being expanded).  A level L is two launches:
//   update_compact(L): for every vertex with next != 0: new = next & ~seen,
//       seen |= new, the level of every source in new recorded, and (degree
//       > 0) an entry (vertex, first CSR entry, edge prefix, mask = new) in
//       the queue of active vertices; a block scan orders the entries of an
//       2048-vertex chunk and one packed atomicAdd per chunk reserves their
//       slots and edge range, as in the single-source compaction.
//   expand(L): edge-parallel tiles over the queue; for edge (u, v) the bits
//       mask(u) & ~seen(v) are OR-ed into next(v) with a 64-bit atomicOr.
// Dense levels post the atomicOr (its result is not needed, so the wave does
// not stall) and the next update_compact scans all n vertices.  Sparse levels
// (few frontier edges) use the returned old value: the edge that turns
// next(v) from 0 to non-zero appends v to a touched list, and the next
// update_compact visits only the listed vertices -- a level then costs
// O(frontier), which is what deep, thin graphs (road networks) need.
// On symmetric graphs a big level runs bottom-up instead: every vertex with
// sources still unseen ORs the seen masks of its neighbours (fmask, a copy
// of seen that update_compact stores when a vertex gains bits) and stops as
// soon as all its unseen sources are covered; next(v) is then a plain
// store.  A neighbour's seen bits are as good as its level-L bits there: a
// source s unseen at v after level L has dist_s(v) > L, so every neighbour
// u has dist_s(u) >= L, and s in seen(u) means dist_s(u) = L.  A level goes
// bottom-up when its frontier edges * alpha exceed the edges of the vertices
// not yet reached by every source.
// Level writes are staged as one byte per (vertex, source): each vertex owns
// a 64-byte line.  The first level that reaches the vertex writes the whole
// line (its new sources' bytes, 0xFF elsewhere) and sets the vertex's bit in
// a `reached` bitmap; later levels read the line, merge the bytes of their
// new sources and write it back whole.  The lines are never cleared: after the last
// level ms_finish transposes the lines of reached vertices into the k rows of
// n int32 (0xFF -> -1) with full 256-byte stores, and writes -1 for the
// others.  Levels from 254 up are stored into the rows
// directly (the wave walks the union of its lanes' new bits) and marked 254
// in the line, which the transpose leaves alone.
// Sources, k and the output buffer are read from device memory (MsParams), so
// the whole sequence (parameter upload, clears, levels) is captured once into
// a hipGraph per kernel shape and replayed for every batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "device_graph.hpp"
#include "frontier.hpp"
#include "wave.hpp"

namespace gg {

namespace {

// ctl[] layout (64-bit words).  Level L's queue counter is cnt[L & 1]
// (count << 32 | edges); tcnt[L & 1] is the length of the touched list that
// feeds update_compact(L) and tvalid[L & 1] says whether that list is in use.
enum : int {
    MS_CNT0 = 0,
    MS_DONE = 2,
    MS_ERR = 3,
    MS_MAXLEV = 4,
    MS_TCNT0 = 5,
    MS_TVALID0 = 7,
    MS_FULLDEG = 9,    // CSR entries of the vertices every source has reached
    MS_BULEVELS = 10,  // levels run bottom-up
    MS_ACT0 = 11,      // act[L & 1]: sources with a non-empty level-L frontier
    MS_DCNT0 = 13,     // 3 words, fused (deep) levels: length of level L's touched list at dcnt[L % 3]
    MS_MAXF = 16,      // fused levels: longest touched list
    MS_MAXE = 17,      // largest frontier (CSR entries) of a top-down level
    MS_BUFIRST = 18,   // first / last level the direction rule chose bottom-up (~0 / 0: none)
    MS_BULAST = 19,
    MS_FMP0 = 20,      // fmp[L & 1]: which mask copy is current at level L (0: fmask, 1: fmask2)
    MS_POST0 = 22,     // post[L & 1]: expand(L-1) updated the light vertices itself (update_compact(L)
                       // takes the heavy ones only)
    MS_WORDS = 24
};
enum : unsigned long long { MS_ERR_TILE = 1, MS_ERR_RANGE = 2 };

// update_compact: 2048-vertex chunks of 512 threads x 4.  Swept on 64-source
// batches: at R-MAT s24, 1024 x 2 took 8.14 ms against 8.22 ms (1024 x 1
// 8.59, 512 x 2 8.69, 512 x 8 8.55, 256 x 8 8.86, 512 x 1 10.3), but on a
// 2000 x 2000 grid (thousands of thin levels) 1024 x 2 took 265 ms against
// 244 ms, and at R-MAT s20 0.88 ms against 0.85 ms.
constexpr int kUcBlock = 512;
constexpr int kUcRounds = 4;
constexpr int kUcChunk = kUcBlock * kUcRounds;
constexpr int kMaxCapturedLevels = 32;  // deeper searches continue with direct launches
constexpr int kByteLevels = 254;         // levels kept in the byte lines; 254 = stored directly, 255 = none
constexpr int kNibLevels = 14;           // levels kept in the nibble lines; 14 = stored directly, 15 = none
constexpr int kMaxLevelBatch = 128;     // levels enqueued between done checks
constexpr int64_t kHubLdsMinVertices = int64_t(1) << 20;  // bottom-up hub masks in LDS from this many vertices

// Per vertex, side by side: a top-down probe reads both in one 16-byte load
// (one cache line instead of two).
struct __align__(16) SeenNext {
    unsigned long long seen;  // sources that reached the vertex
    unsigned long long next;  // sources reaching it in the level being expanded
};

struct MsArgs {
    const int64_t* __restrict__ offsets;
    const int32_t* __restrict__ cols;
    SeenNext* sn;
    unsigned long long* fmask;   // per vertex: seen, stored when it changes (the mask bottom-up levels gather)
    unsigned long long* fmask2;  // the other copy: a bottom-up level that updates its vertices itself
                                 // writes the new seen of every vertex there and gathers from the first
    unsigned long long* lv8;     // per vertex: its level line (8 words of bytes, or 4 of nibbles when nib)
    unsigned long long* reached; // bit per vertex: its lv8 line has been written in this batch
    MsParams* p;                 // sources, k, output rows (device memory)
    const MsParams* hp;          // the same, host-visible (written by the host before the launch)
    const int32_t* perm;         // original id -> id of the degree-ordered copy
    const int32_t* iperm;        // id of the degree-ordered copy -> original id
    const uint32_t* heavy_tiles; // rounds + 1: first tile of each round of the heavy rows
    int64_t heavy;               // vertices [0, heavy) have rows of >= kMsHeavyRow entries
    int32_t heavy_rounds;
    int64_t nz;                  // vertices [nz, n) have no edges: reached only as sources, at level 0
    uint32_t nz_chunks;          // update chunks covering [0, nz)
    int64_t* q_rb;
    uint32_t* q_eoff;
    int32_t* q_v;
    unsigned long long* q_fm;    // frontier mask of each queue entry
    uint32_t* tile_start;
    int32_t* touched;
    int32_t* touched2;           // fused levels: touched list of odd levels (even ones use touched)
    unsigned long long* next1;   // fused levels: next of odd levels (even ones use sn's next)
    unsigned long long* ctl;
    int64_t n;
    int64_t m;
    uint32_t nchunks;
    uint32_t tile_edges;
    uint32_t sparse_edges;
    uint32_t bottom_up;          // bottom-up levels allowed (symmetric graph)
    uint32_t alpha;
    uint32_t low_degree;         // bottom-up: longer rows are scanned by the whole wave
    uint32_t beta;               // bottom-up only while F * beta >= nz (0: no such limit)
    uint32_t nib;                // level lines of nibbles (shallow searches), else of bytes
    uint32_t hub_lds;            // bottom-up levels keep the masks of the first vertices in LDS
};

template <int BLOCK, int K>
struct MsTileLds {
    static constexpr int T = BLOCK * K;
    uint32_t eoff[T + 2];
    int64_t rb[T + 1];
    unsigned long long fm[T + 1];
};

// Bottom-up levels: the masks of the first kHubPerThread * BLOCK vertices
// (the highest degrees: at R-MAT s24 the first 10,240 of 16.8 M vertices take
// about 22% of all row entries) are copied into LDS at the start of the
// level, so that those gathers do not go to the L2.  10 KB per wave: with 16
// waves per CU that fills its 160 KB.
constexpr int kHubPerThread = 20;

template <int BLOCK, int K>
union MsExpandLds {
    MsTileLds<BLOCK, K> td;
    unsigned long long hub[BLOCK * kHubPerThread];
};

// seen = next = 0; no vertex reached (the level lines are left as they are).
// Vertices without edges (ids from nz on, symmetric graphs) are only ever
// touched as sources, so after the first batch only [0, nz) is cleared here
// and ms_seed clears the previous batch's sources.
__global__ void ms_clear_kernel(MsArgs a) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t tid0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    // The batch's parameters come from the host-visible block; workgroup 0
    // makes the device copy the later kernels read.
    static_assert(sizeof(MsParams) % 4 == 0, "parameter block copied in words");
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < sizeof(MsParams) / 4; i += blockDim.x)
            reinterpret_cast<uint32_t*>(a.p)[i] = reinterpret_cast<const uint32_t*>(a.hp)[i];
    const int64_t end = a.hp->clear_all ? a.n : a.nz;
    for (int64_t i = tid0; i < end; i += stride) {
        a.sn[i].seen = 0;
        a.sn[i].next = 0;
        a.next1[i] = 0;
        a.fmask[i] = 0;
    }
    for (int64_t i = tid0; i < (a.n + 63) / 64; i += stride) a.reached[i] = 0;
}

// Byte i of the result is 0xFF when bit i of b is set (b < 256).
__device__ __forceinline__ unsigned long long spread_bits_to_bytes(unsigned long long b) {
    const unsigned long long m = (b * 0x0101010101010101ull) & 0x8040201008040201ull;
    const unsigned long long nz = (((m & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | m) & 0x8080808080808080ull;
    return (nz >> 7) * 0xFFull;
}

// Nibble i of the result is 0xF when bit i of b (16 bits) is set.
__device__ __forceinline__ unsigned long long spread_bits_to_nibbles(unsigned long long b) {
    unsigned long long x = b & 0xFFFFull;
    x = (x | (x << 24)) & 0x000000FF000000FFull;
    x = (x | (x << 12)) & 0x000F000F000F000Full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
    return x * 0xFull;
}

// Level lines: per vertex, the level at which each source reached it.  Byte
// lines (8 words) hold byte s for source s; nibble lines (4 words, for
// searches known to stay shallow) nibble s, which halves the bytes every
// level writes and the finish reads.  A format's largest value means "not
// reached", the one below "stored in the output rows directly" (levels
// from kByteLevels or kNibLevels on).
__device__ __forceinline__ int line_pairs(const MsArgs& a) { return a.nib ? 2 : 4; }  // 16-byte pairs
__device__ __forceinline__ int line_levels(const MsArgs& a) { return a.nib ? kNibLevels : kByteLevels; }
__device__ __forceinline__ const ulonglong2* line_of(const MsArgs& a, int64_t v) {
    return reinterpret_cast<const ulonglong2*>(a.lv8 + v * (a.nib ? 4 : 8));
}

// After the last level: level rows [k][n] from the level lines, a wave per 64
// vertices, one full 256-byte store per source (a line's "not reached" value
// -> -1; levels stored directly are left alone).
// `snap` (host-visible, may be null): where workgroup 0 stores the control
// block, which no kernel changes from here on.
template <bool NIB>
__global__ __launch_bounds__(256) void ms_finish_kernel(MsArgs a, int only_if_done, unsigned long long* snap) {
    if (snap && blockIdx.x == 0 && threadIdx.x < Runner::kCtlWords) snap[threadIdx.x] = a.ctl[threadIdx.x];
    // Replayed right after the first levels: a no-op unless the search has ended.
    if (only_if_done && (!a.ctl[MS_DONE] || a.ctl[MS_ERR])) return;
    const int k = a.p->k;
    int32_t* const levels = a.p->levels;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t ngroups = (static_cast<uint64_t>(a.n) + 63) / 64;
    const uint64_t wstride = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); g < ngroups;
         g += wstride) {
        const int64_t v = static_cast<int64_t>(g * 64 + lane);  // original id: the output position
        const bool valid = v < a.n;
        unsigned long long w[8];
        const int64_t vn = valid ? a.perm[v] : 0;
        const bool reached = valid && ((a.reached[vn >> 6] >> (vn & 63)) & 1ull);
        constexpr int kPairs = NIB ? 2 : 4;
        const ulonglong2* line = reinterpret_cast<const ulonglong2*>(a.lv8 + vn * 2 * kPairs);
#pragma unroll
        for (int j = 0; j < kPairs; ++j) {
            const ulonglong2 x = reached ? line[j] : make_ulonglong2(~0ull, ~0ull);
            w[2 * j] = x.x;
            w[2 * j + 1] = x.y;
        }
        constexpr uint32_t kBits = NIB ? 4 : 8, kNone = (1u << kBits) - 1;
        constexpr uint32_t kDirect = NIB ? kNibLevels : kByteLevels;
        // The rows are global memory: say so, so that the stores are global
        // (not flat) stores.
        __attribute__((address_space(1))) int32_t* const rows =
            (__attribute__((address_space(1))) int32_t*)levels;
#pragma unroll
        for (int s = 0; s < 64; ++s) {
            if (s < k) {
                const uint32_t x = static_cast<uint32_t>(w[s / (64 / kBits)] >> (kBits * (s % (64 / kBits)))) & kNone;
                if (valid && x != kDirect)  // output only: streamed past the caches
                    __builtin_nontemporal_store(x == kNone ? -1 : static_cast<int32_t>(x), &rows[int64_t(s) * a.n + v]);
            }
        }
    }
}

// The control block to host-visible memory between batches of levels (the
// earlier kernels' updates are complete at its start).
__global__ __launch_bounds__(64) void ms_snapshot_kernel(const unsigned long long* ctl, unsigned long long* out) {
    if (threadIdx.x < Runner::kCtlWords) out[threadIdx.x] = ctl[threadIdx.x];
}

// Level 0: next(source i) |= bit i; each distinct source once on the touched
// list of level 0.
__global__ void ms_seed_kernel(MsArgs a) {
    if (threadIdx.x == 0) {
        for (int i = 0; i < MS_WORDS; ++i) a.ctl[i] = 0;
        a.ctl[MS_BUFIRST] = ~0ull;
    }
    if (!a.p->clear_all && threadIdx.x < static_cast<uint32_t>(a.p->prev_k)) {
        const int32_t s = a.perm[a.p->prev_sources[threadIdx.x]];
        if (s >= a.nz) {
            a.sn[s].seen = 0;
            a.sn[s].next = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x < static_cast<uint32_t>(a.p->k)) {
        const int32_t s = a.perm[a.p->sources[threadIdx.x]];
        const unsigned long long old = atomicOr(&a.sn[s].next, 1ull << threadIdx.x);
        if (old == 0) a.touched[atomicAdd(&a.ctl[MS_TCNT0], 1ull)] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.ctl[MS_TVALID0] = 1;
        a.ctl[MS_DCNT0] = a.ctl[MS_TCNT0];  // the list of level 0 for fused levels
    }
}

// Stores level L for the new bits of this lane's vertex: the wave walks the
// union of its lanes' bits, one predicated store per source present.
__device__ __forceinline__ void write_levels(int32_t* levels, int64_t n, int64_t v, unsigned long long nv, int L) {
    unsigned long long any = nv;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) any |= __shfl_xor(any, o, 64);
    while (any) {
        const int s = __ffsll(static_cast<long long>(any)) - 1;
        any &= any - 1;
        if ((nv >> s) & 1ull) levels[int64_t(s) * n + v] = L;
    }
}

// Word of a level line with byte s set to the byte of `vb` for every bit s
// of the 8-bit group nv8, other bytes kept from `old`.
__device__ __forceinline__ unsigned long long merge_level_word(unsigned long long old, unsigned long long nv8,
                                                               unsigned long long vb) {
    const unsigned long long bm = spread_bits_to_bytes(nv8);
    return (old & ~bm) | (vb & bm);
}

// Pair j of a level line with the entries of the sources in nv set to level
// value lv (already clamped to the format), other entries kept from `old`.
__device__ __forceinline__ ulonglong2 merge_line_pair(const MsArgs& a, ulonglong2 old, unsigned long long nv, int j,
                                                      unsigned long long lv) {
    if (a.nib) {
        const unsigned long long vn = lv * 0x1111111111111111ull;
        const unsigned long long m0 = spread_bits_to_nibbles(nv >> (32 * j));
        const unsigned long long m1 = spread_bits_to_nibbles(nv >> (32 * j + 16));
        return make_ulonglong2((old.x & ~m0) | (vn & m0), (old.y & ~m1) | (vn & m1));
    }
    const unsigned long long vb = lv * 0x0101010101010101ull;
    return make_ulonglong2(merge_level_word(old.x, (nv >> (16 * j)) & 0xFFull, vb),
                           merge_level_word(old.y, (nv >> (16 * j + 8)) & 0xFFull, vb));
}

// Chunk c of update_compact(L): thread t takes slots c*span + r*kUcBlock + t
// (r < kUcRounds, below (c+1)*span) of either the vertex range (scan) or the
// touched list.  span = kUcChunk, or kUcBlock for a short list, which then
// spreads over all workgroups instead of a few.
//
// The loads of all of a thread's slots are issued before any result is used
// and before any store (slot vertices; then state and row bounds; then the
// level lines that already hold bytes), so a chunk costs three dependent
// memory round trips, not three per slot.  A vertex's line is rewritten
// whole (four 16-byte stores): its first level writes 0xFF for the sources
// that have not reached it, later levels merge their bytes into the line
// they read.  Each vertex is in at most one slot of a launch (the touched
// list holds distinct vertices), so no other thread writes its line or state.
__device__ void ms_uc_chunk(const MsArgs& a, int L, uint32_t c, int64_t span, bool list, uint32_t count,
                            unsigned long long kmask, bool no_queue, unsigned long long* fm, bool post,
                            int64_t scan_end, unsigned long long& fulldeg, unsigned long long& act,
                            unsigned long long& wg_count) {
    __shared__ unsigned long long s_wave[kUcBlock / 64];
    __shared__ unsigned long long s_excl;
    __shared__ uint32_t s_ok;
    const uint32_t tid = threadIdx.x;
    const int64_t ibase = int64_t(c) * span;
    const int64_t iend = ibase + span;
    int32_t* levels = a.p->levels;
    // Step 1: the vertex of each slot.
    int64_t vs[kUcRounds];
#pragma unroll
    for (int r = 0; r < kUcRounds; ++r) {
        const int64_t i = ibase + r * kUcBlock + tid;
        int64_t v = -1;
        if (list) {
            if (i < count && i < iend) v = a.touched[i];
        } else if (i < scan_end) {  // a scan comes after level 0: vertices without edges get nothing
            v = i;
        }
        vs[r] = v;
    }
    // Step 2: seen / next and the row bounds.
    ulonglong2 st[kUcRounds];
    int64_t rb[kUcRounds], re[kUcRounds];
#pragma unroll
    for (int r = 0; r < kUcRounds; ++r) {
        st[r] = make_ulonglong2(0ull, 0ull);
        rb[r] = re[r] = 0;
        if (vs[r] >= 0) {
            st[r] = *reinterpret_cast<const ulonglong2*>(a.sn + vs[r]);
            rb[r] = a.offsets[vs[r]];
            re[r] = a.offsets[vs[r] + 1];
        }
    }
    // Step 3: new bits; the lines of vertices reached at an earlier level.
    unsigned long long newbits[kUcRounds];
    bool first[kUcRounds];
    ulonglong2 line[kUcRounds][4];
#pragma unroll
    for (int r = 0; r < kUcRounds; ++r) {
        const unsigned long long sv = st[r].x;
        newbits[r] = st[r].y & ~sv;
        first[r] = newbits[r] && !sv;  // the first level that reaches the vertex
        const ulonglong2* src = line_of(a, vs[r] >= 0 ? vs[r] : 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            line[r][j] = (newbits[r] && !first[r] && j < line_pairs(a)) ? src[j] : make_ulonglong2(~0ull, ~0ull);
    }
    // Step 4: stores.
    uint32_t active = 0;  // bit r: the vertex of slot r goes to the queue
    uint32_t kept = 0, edges = 0;
    const unsigned long long lv = static_cast<unsigned long long>(L < line_levels(a) ? L : line_levels(a));
#pragma unroll
    for (int r = 0; r < kUcRounds; ++r) {
        const int64_t v = vs[r];
        const unsigned long long nv = newbits[r];
        if (v >= 0) {
            if (st[r].y) *reinterpret_cast<ulonglong2*>(a.sn + v) = make_ulonglong2(st[r].x | nv, 0ull);
            // seen only changes here; after a level that updated the light
            // vertices itself, every heavy vertex's copy is written (the
            // copy in use changed)
            if (nv || post) fm[v] = st[r].x | nv;
        }
        act |= nv;
        if (list) {
            if (first[r]) atomicOr(&a.reached[v >> 6], 1ull << (v & 63));
        } else {
            // a wave holds the 64 vertices of one bitmap word, which no other
            // wave of this kernel touches
            const uint64_t fb = wave_ballot(first[r]);
            const int64_t w = (ibase + r * kUcBlock + (tid & ~63u)) >> 6;
            if (fb && (tid & 63u) == 0) a.reached[w] |= fb;
        }
        if (L >= line_levels(a)) {
            // every lane of the wave takes part in the shuffles of write_levels
            if (__ballot(nv != 0)) write_levels(levels, a.n, nv ? a.iperm[v] : 0, nv, L);
        }
        if (nv) {
            ulonglong2* dst = const_cast<ulonglong2*>(line_of(a, v));
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < line_pairs(a)) dst[j] = merge_line_pair(a, line[r][j], nv, j, lv);
            const uint32_t d = static_cast<uint32_t>(re[r] - rb[r]);
            if (d) {
                active |= 1u << r;
                ++kept;
                edges += d;
            }
            if (((st[r].x | nv) & kmask) == kmask) fulldeg += d;  // every source has reached v now
        }
    }
    bool anynew = false;
#pragma unroll
    for (int r = 0; r < kUcRounds; ++r) anynew |= newbits[r] != 0;
    anynew = __syncthreads_or(anynew);
    const unsigned long long mine = (static_cast<unsigned long long>(kept) << 32) | edges;
    unsigned long long total;
    const unsigned long long excl_local = block_exclusive_scan<kUcBlock>(mine, s_wave, total);
    if (tid == 0) {
        s_ok = 1;
        if (total && no_queue) {
            // no slots to reserve: the counts go to ctl once per workgroup
            // (an atomic per chunk on the one counter serialises)
            wg_count += total;
        } else if (total) {
            s_excl = atomicAdd(&a.ctl[MS_CNT0 + (L & 1)], total);
            if (!reservation_fits(s_excl, total, a.n, a.m)) {
                atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
                s_ok = 0;
            }
        }
        // Every writer of this level stores the same L and levels run in
        // order, so the final value is the deepest level reached.
        if (anynew) a.ctl[MS_MAXLEV] = static_cast<unsigned long long>(L);
    }
    __syncthreads();
    // A level that goes bottom-up for sure reads no queue: only its counts.
    if (!active || !s_ok || no_queue) return;
    const unsigned long long pre = s_excl + excl_local;
    uint32_t slot = static_cast<uint32_t>(pre >> 32);
    uint32_t eo = static_cast<uint32_t>(pre);
#pragma unroll
    for (int r = 0; r < kUcRounds; ++r) {
        if ((active >> r) & 1u) {
            const uint32_t d = static_cast<uint32_t>(re[r] - rb[r]);
            a.q_rb[slot] = rb[r];
            a.q_eoff[slot] = eo;
            a.q_v[slot] = static_cast<int32_t>(vs[r]);
            a.q_fm[slot] = newbits[r];
            mark_tiles(a.tile_start, a.tile_edges, eo, d, slot);
            ++slot;
            eo += d;
        }
    }
}

// Level L, step 1: assign level L to the bits found by expand(L-1) and queue
// the vertices that carry them.
__global__ __launch_bounds__(kUcBlock) void ms_update_compact_kernel(MsArgs a, int L) {
    if (a.ctl[MS_DONE] | a.ctl[MS_ERR]) return;
    const int par = L & 1;
    const bool list = a.ctl[MS_TVALID0 + par] != 0;
    const unsigned long long count = list ? a.ctl[MS_TCNT0 + par] : 0;
    // cnt[(L+1) & 1] was last read by expand(L-1) and tcnt[(L+1) & 1] by
    // update_compact(L-1): both restart from zero for level L+1.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl[MS_CNT0 + ((L + 1) & 1)] = 0;
        a.ctl[MS_TCNT0 + ((L + 1) & 1)] = 0;
        a.ctl[MS_ACT0 + ((L + 1) & 1)] = 0;  // last read by expand(L-1)
    }
    if (count > static_cast<unsigned long long>(a.n)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
        return;
    }
    const int64_t span = list && count <= static_cast<unsigned long long>(kUcChunk) * gridDim.x ? kUcBlock : kUcChunk;
    // After a bottom-up level that updated the light vertices itself, only
    // the heavy prefix [0, heavy) is left.
    const bool post = !list && a.ctl[MS_POST0 + par] != 0;
    unsigned long long* const fm = a.ctl[MS_FMP0 + par] ? a.fmask2 : a.fmask;
    const int64_t scan_end = post ? a.heavy : a.nz;
    const uint32_t nch = list   ? static_cast<uint32_t>((count + span - 1) / span)
                         : post ? static_cast<uint32_t>((a.heavy + kUcChunk - 1) / kUcChunk)
                                : a.nz_chunks;
    const int k = a.p->k;
    const unsigned long long kmask = k >= 64 ? ~0ull : ((1ull << k) - 1);
    // Levels in the learned bottom-up window go bottom-up after a scan
    // (expand() forces it), so their queue is not built.
    const bool no_queue = a.bottom_up && !list && L >= a.p->bu_lo && L <= a.p->bu_hi;
    unsigned long long fulldeg = 0, act = 0, wg_count = 0;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        ms_uc_chunk(a, L, c, span, list, static_cast<uint32_t>(count), kmask, no_queue, fm, post, scan_end, fulldeg,
                    act, wg_count);
        __syncthreads();
    }
    // One atomic per workgroup for the edges of the newly fully seen
    // vertices, one for the sources that reached new vertices, and (levels
    // without a queue) one for the level's counts.
    __shared__ unsigned long long s_fd[kUcBlock / 64], s_act[kUcBlock / 64];
    const unsigned long long w = wave_sum(fulldeg);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) act |= __shfl_xor(act, o, 64);
    if ((threadIdx.x & 63) == 0) {
        s_fd[threadIdx.x >> 6] = w;
        s_act[threadIdx.x >> 6] = act;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0, b = 0;
        for (int i = 0; i < kUcBlock / 64; ++i) {
            t += s_fd[i];
            b |= s_act[i];
        }
        if (t) atomicAdd(&a.ctl[MS_FULLDEG], t);
        if (b) atomicOr(&a.ctl[MS_ACT0 + (L & 1)], b);
        if (wg_count) atomicAdd(&a.ctl[MS_CNT0 + (L & 1)], wg_count);
    }
}

// Level L bottom-up: next(v) = the unseen sources of v found among the
// frontier masks of its neighbours; `kmask` = the sources still searching.
// Heavy rows go in tiles over all workgroups (below).  Of the other rows, a
// wave owns 64 consecutive vertices; rows of at most `low_degree` entries
// are scanned by their lane, longer rows by the whole wave 64 entries at a
// time; a scan stops once all of the vertex's unseen sources are covered.
constexpr int kSubStep = 4;     // bottom-up long rows: entries per lane in a sub-wave's later steps
constexpr int kHeavyBatch = 4;  // bottom-up heavy rows: tiles a wave scans together

// Per-thread totals of the light vertices a bottom-up level updates itself.
struct InlineTotals {
    unsigned long long cnt = 0;      // (vertices << 32) | edges queued for level L+1
    unsigned long long fulldeg = 0;  // edges of the vertices every source has now reached
    unsigned long long act = 0;      // sources that reached new vertices
    bool anynew = false;
};

// The update of update_compact(L+1) for light vertex v (lane of the wave that
// owns the 64 vertices of reached word g; every lane calls it): seen and its
// new copy in fm_out, level L+1 into the line of the sources in nb, the
// reached bit, and the totals.  No queue entry: level L+1 goes bottom-up.
__device__ __forceinline__ void ms_inline_update(const MsArgs& a, unsigned long long* fm_out, int64_t v, uint64_t g,
                                                 bool mine, unsigned long long sv, unsigned long long nb,
                                                 uint32_t deg, int L1, unsigned long long kfull, InlineTotals& t) {
    const bool first = nb && !sv;
    const ulonglong2* src = line_of(a, mine ? v : 0);
    ulonglong2 line[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        line[j] = (nb && !first && j < line_pairs(a)) ? src[j] : make_ulonglong2(~0ull, ~0ull);
    if (mine) {
        fm_out[v] = sv | nb;
        if (nb) *reinterpret_cast<ulonglong2*>(a.sn + v) = make_ulonglong2(sv | nb, 0ull);
    }
    // this wave alone writes word g in this kernel (the heavy vertices of
    // the word are updated by the next update_compact)
    const uint64_t fb = wave_ballot(first);
    if (fb && (threadIdx.x & 63u) == 0) a.reached[g] |= fb;
    if (L1 >= line_levels(a)) {
        // every lane of the wave takes part in the shuffles of write_levels
        if (__ballot(nb != 0)) write_levels(a.p->levels, a.n, nb ? a.iperm[v] : 0, nb, L1);
    }
    if (nb) {
        const unsigned long long lv = static_cast<unsigned long long>(L1 < line_levels(a) ? L1 : line_levels(a));
        ulonglong2* dst = const_cast<ulonglong2*>(line_of(a, v));
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < line_pairs(a)) dst[j] = merge_line_pair(a, line[j], nb, j, lv);
        t.cnt += (1ull << 32) | deg;
        if (((sv | nb) & kfull) == kfull) t.fulldeg += deg;
        t.act |= nb;
        t.anynew = true;
    }
}

// fm_in: the mask copy current at level L.  fm_out (null: the level only
// stores next): the light vertices are updated here for level L1 = L + 1
// (ms_inline_update) and every light vertex's seen goes to fm_out; kfull =
// all k sources.  `hub` (LDS, at least 64 words) also holds the workgroup's
// totals at the end.
__device__ void ms_bottom_up(const MsArgs& a, unsigned long long kmask, unsigned long long* hub, uint32_t nhub,
                             const unsigned long long* fm_in, unsigned long long* fm_out, int L1,
                             unsigned long long kfull) {
    const uint32_t lane = threadIdx.x & 63u;
    bool bad = false;
    InlineTotals tot;
    // Vertex u's mask (u in [0, n)): from the LDS copy for the first nhub.
    auto mask_of = [&](int32_t u) { return static_cast<uint32_t>(u) < nhub ? hub[u] : fm_in[u]; };
    // Heavy rows (a prefix of the degree-ordered ids): tiles of kMsHeavyRow
    // entries spread over every wave, so the longest rows do not hold up a
    // few waves.  Tiles are numbered round by round (tile r of every row
    // that has one), so a row's first tile is scanned early and its later
    // tiles are skipped once the row's unseen sources are covered.  A wave
    // checks 64 tiles at once (a lane each), then scans the open ones
    // together, kMsHeavyRow / 64 entries per lane, and ORs what it finds
    // into next.
    const uint32_t ntiles = a.heavy_rounds > 0 ? a.heavy_tiles[a.heavy_rounds] : 0u;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    for (uint32_t t0 = (blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * 64u; t0 < ntiles;
         t0 += nwaves * 64u) {
        const uint32_t t = t0 + lane;
        int64_t v = 0, rb = 0, re = 0;
        unsigned long long un = 0;
        if (t < ntiles) {
            int32_t lo = 0, hi = a.heavy_rounds - 1;  // round of tile t: last round starting at or before t
            while (lo < hi) {
                const int32_t mid = (lo + hi + 1) / 2;
                if (a.heavy_tiles[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            v = static_cast<int64_t>(t - a.heavy_tiles[lo]);  // the rows of a round are a prefix
            const unsigned long long got =
                __hip_atomic_load(&a.sn[v].next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            un = kmask & ~a.sn[v].seen;
            if (!(un & ~got)) un = 0;
            rb = a.offsets[v] + int64_t(lo) * kMsHeavyRow;
            re = std::min<int64_t>(a.offsets[v + 1], rb + kMsHeavyRow);
        }
        // The open tiles, kHeavyBatch at a time: the loads of all of them
        // are issued before any is used.
        uint64_t om = wave_ballot(un != 0);
        while (om) {
            int64_t hv[kHeavyBatch], hrb[kHeavyBatch], hre[kHeavyBatch];
            unsigned long long hun[kHeavyBatch];
#pragma unroll
            for (int g = 0; g < kHeavyBatch; ++g) {
                const bool use = om != 0;  // wave-uniform
                const int hl = use ? __ffsll(static_cast<long long>(om)) - 1 : 0;
                om &= om - 1;
                hv[g] = __shfl(v, hl, 64);
                hrb[g] = __shfl(rb, hl, 64);
                hre[g] = use ? __shfl(re, hl, 64) : hrb[g];
                hun[g] = use ? __shfl(un, hl, 64) : 0ull;
            }
            int32_t u[kHeavyBatch][kMsHeavyRow / 64];
#pragma unroll
            for (int g = 0; g < kHeavyBatch; ++g)
#pragma unroll
                for (int j = 0; j < kMsHeavyRow / 64; ++j) {
                    const int64_t i = hrb[g] + j * 64 + lane;
                    u[g][j] = i < hre[g] ? a.cols[i] : -1;
                }
            unsigned long long acc[kHeavyBatch];
#pragma unroll
            for (int g = 0; g < kHeavyBatch; ++g) {
                acc[g] = 0;
#pragma unroll
                for (int j = 0; j < kMsHeavyRow / 64; ++j) {
                    if (u[g][j] < 0) continue;
                    if (u[g][j] >= a.n) bad = true;
                    else acc[g] |= mask_of(u[g][j]);
                }
            }
#pragma unroll
            for (int g = 0; g < kHeavyBatch; ++g) {
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) acc[g] |= __shfl_xor(acc[g], o, 64);
                if (lane == 0 && (acc[g] & hun[g])) atomicOr(&a.sn[hv[g]].next, acc[g] & hun[g]);
            }
        }
    }
    // Light rows: a wave owns 64 consecutive vertices (vertices without edges,
    // the tail of the order, find nothing and are left out).
    const uint64_t ngroups = (static_cast<uint64_t>(a.nz) + 63) / 64;
    const uint64_t wstride = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
    const uint64_t g_first = static_cast<uint64_t>(a.heavy) / 64;
    for (uint64_t g = g_first + static_cast<uint64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
         g < ngroups; g += wstride) {
        const int64_t v = static_cast<int64_t>(g * 64 + lane);
        unsigned long long un = 0, sv = 0;
        const bool mine = v >= a.heavy && v < a.nz;
        if (mine) {
            sv = a.sn[v].seen;
            un = kmask & ~sv;
        }
        if (!wave_ballot(un != 0)) {
            if (fm_out && mine) fm_out[v] = sv;
            continue;
        }
        int64_t rb = 0, re = 0;
        if (un) {
            rb = a.offsets[v];
            re = a.offsets[v + 1];
        }
        const uint32_t deg = static_cast<uint32_t>(re - rb);
        unsigned long long acc = 0;
        if (un && deg <= a.low_degree) {
            for (int64_t i = rb; i < re; i += 4) {
                int32_t u[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) u[j] = (i + j < re) ? a.cols[i + j] : -1;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (u[j] < 0) continue;
                    if (u[j] >= a.n) {
                        bad = true;
                        continue;
                    }
                    acc |= mask_of(u[j]);
                }
                if ((acc & un) == un) break;
            }
        }
        // Longer rows: 8 at a time, one per 8-lane sub-wave, 8 entries per
        // step; each sub-wave stops when its row's unseen sources are
        // covered, and the wave goes on while any sub-wave has work.
        uint64_t hm = wave_ballot(un && deg > a.low_degree);
        const uint32_t sg = lane >> 3, sl = lane & 7u;
        while (hm) {
            uint64_t mm = hm;  // this sub-wave's row: the sg-th set bit of hm
            for (uint32_t k = 0; k < sg; ++k) mm &= mm - 1;
            const int hl = mm ? __ffsll(static_cast<long long>(mm)) - 1 : static_cast<int>(lane);
            // (every lane takes part in the shuffles: a lane reads only from
            // active lanes)
            const int64_t hrb = __shfl(rb, hl, 64);
            const int64_t hre_row = __shfl(re, hl, 64);
            const int64_t hre = mm ? hre_row : hrb;
            const unsigned long long hun = __shfl(un, hl, 64);
            unsigned long long hacc = 0;
            bool open = hrb < hre;
            // The first step takes 8 entries (a row often closes on its first
            // hubs); a row still open after it is usually scanned to its end,
            // so later steps take kSubStep entries per lane, all loads issued
            // before any is used.
            int per = 1;
            for (int64_t b0 = hrb; wave_ballot(open); b0 += 8 * per, per = kSubStep) {
                int32_t u[kSubStep];
                bool ok[kSubStep];
#pragma unroll
                for (int j = 0; j < kSubStep; ++j) {
                    const int64_t i = b0 + j * 8 + sl;
                    ok[j] = j < per && open && i < hre;
                    u[j] = ok[j] ? a.cols[i] : 0;
                }
                unsigned long long x = 0;
#pragma unroll
                for (int j = 0; j < kSubStep; ++j) {
                    if (!ok[j]) continue;
                    if (u[j] < 0 || u[j] >= a.n) bad = true;
                    else x |= mask_of(u[j]);
                }
#pragma unroll
                for (int o = 4; o > 0; o >>= 1) x |= __shfl_xor(x, o, 8);
                if (open) {
                    hacc |= x;
                    open = (hacc & hun) != hun && b0 + 8 * per < hre;  // same in the 8 lanes of the sub-wave
                }
            }
            // The owner of each row takes its sub-wave's result.
            const uint64_t round = hm;
            for (uint32_t k = 0; k < 8 && hm; ++k) hm &= hm - 1;  // the rows taken this round
            const uint64_t taken = round & ~hm;
            const uint32_t rank = static_cast<uint32_t>(__popcll(round & ((1ull << lane) - 1)));
            const unsigned long long got = __shfl(hacc, static_cast<int>((rank & 7u) * 8), 64);
            if ((taken >> lane) & 1ull) acc = got;
        }
        const unsigned long long nb = acc & un;
        if (fm_out) ms_inline_update(a, fm_out, v, g, mine, sv, nb, deg, L1, kfull, tot);
        else if (nb) a.sn[v].next = nb;
    }
    if (bad) atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
    if (!fm_out) return;
    // One atomic per workgroup for each total of level L1, summed in the
    // hub words once every wave is past its last read of them.
    const unsigned long long c = wave_sum(tot.cnt), f = wave_sum(tot.fulldeg);
    unsigned long long b = tot.act;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) b |= __shfl_xor(b, o, 64);
    const bool anynew = __ballot(tot.anynew) != 0;
    const uint32_t w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        hub[w] = c;
        hub[16 + w] = f;
        hub[32 + w] = b;
        hub[48 + w] = anynew ? 1ull : 0ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tc = 0, tf = 0, tb = 0, tn = 0;
        for (uint32_t i = 0; i < blockDim.x / 64; ++i) {
            tc += hub[i];
            tf += hub[16 + i];
            tb |= hub[32 + i];
            tn |= hub[48 + i];
        }
        if (tc) atomicAdd(&a.ctl[MS_CNT0 + (L1 & 1)], tc);
        if (tf) atomicAdd(&a.ctl[MS_FULLDEG], tf);
        if (tb) atomicOr(&a.ctl[MS_ACT0 + (L1 & 1)], tb);
        // every writer of level L1 stores the same value
        if (tn) a.ctl[MS_MAXLEV] = static_cast<unsigned long long>(L1);
    }
}

// Level L, step 2: push the queue's masks along every edge.
template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void ms_expand_kernel(MsArgs a, int L) {
    __shared__ MsExpandLds<BLOCK, K> lds;
    MsTileLds<BLOCK, K>& s = lds.td;
    constexpr uint32_t T = BLOCK * K;
    if (a.ctl[MS_DONE] | a.ctl[MS_ERR]) return;
    const unsigned long long cnt = a.ctl[MS_CNT0 + (L & 1)];
    const uint32_t F = static_cast<uint32_t>(cnt >> 32), E = static_cast<uint32_t>(cnt);
    if (F == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[MS_DONE] = 1;
        return;
    }
    if (F > a.n || E > a.m) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
        return;
    }
    const bool sparse = E <= a.sparse_edges;
    const int npar = (L + 1) & 1;
    // Bottom-up only after a scanning update_compact(L): a level handed a
    // touched list is sparse, and top-down is cheaper there.  MS_FULLDEG is
    // final (update_compact(L) has ended).
    const bool scanned = a.ctl[MS_TVALID0 + (L & 1)] == 0;
    const int64_t open_edges = a.m - static_cast<int64_t>(a.ctl[MS_FULLDEG]);
    // Beamer's second rule as well: once the frontier holds fewer than
    // nz / beta vertices, a top-down level (which hands a list to the next
    // update) is cheaper than a bottom-up one (after which the next update
    // scans every vertex).
    const bool rule = a.bottom_up && scanned && static_cast<int64_t>(E) * a.alpha > open_edges &&
                      (a.beta == 0 || static_cast<int64_t>(F) * a.beta >= a.nz);
    // Levels in the window the rule chose in the runner's last batch go
    // bottom-up whatever it says now: their update built no queue (bottom-up
    // levels give the same levels as top-down ones, only the cost differs).
    // The window learned for the next batch is the rule's own choice.
    const bool forced = a.bottom_up && scanned && L >= a.p->bu_lo && L <= a.p->bu_hi;
    const bool bottom_up = rule || forced;
    // When level L+1 is in the window too (bottom-up after a scan, so no
    // queue), this level updates its light vertices itself, for level L+1:
    // update_compact(L+1) is left the heavy prefix.  The new seen of every
    // vertex then goes to the other mask copy (the neighbours of a vertex
    // updated early must still read its level-L mask), which is current from
    // level L+1 on.
    const bool inline_update = bottom_up && L + 1 >= a.p->bu_lo && L + 1 <= a.p->bu_hi;
    const unsigned long long fmp = a.ctl[MS_FMP0 + (L & 1)];
    const unsigned long long* const fm_in = fmp ? a.fmask2 : a.fmask;
    unsigned long long* const fm_out = fmp ? a.fmask : a.fmask2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl[MS_FMP0 + npar] = inline_update ? (fmp ^ 1ull) : fmp;
        a.ctl[MS_POST0 + npar] = inline_update ? 1ull : 0ull;
        a.ctl[MS_TVALID0 + npar] = (sparse && !bottom_up) ? 1 : 0;
        if (bottom_up) atomicAdd(&a.ctl[MS_BULEVELS], 1ull);
        else atomicMax(&a.ctl[MS_MAXE], static_cast<unsigned long long>(E));
        if (rule) {
            atomicMin(&a.ctl[MS_BUFIRST], static_cast<unsigned long long>(L));
            atomicMax(&a.ctl[MS_BULAST], static_cast<unsigned long long>(L));
        }
    }
    if (bottom_up) {
        // Only sources with a non-empty level-L frontier can reach anything
        // at level L+1: a source that has finished (say, in a small
        // component) would otherwise keep every row it never reaches open.
        const int k = a.p->k;
        const unsigned long long kmask = k >= 64 ? ~0ull : ((1ull << k) - 1);
        // The mask copy does not change during the level (update_compact
        // stores it, or an inline update stores the other copy).
        const uint32_t nhub = a.hub_lds ? static_cast<uint32_t>(std::min<int64_t>(BLOCK * kHubPerThread, a.n)) : 0u;
        for (uint32_t i = threadIdx.x; i < nhub; i += BLOCK) lds.hub[i] = fm_in[i];
        __syncthreads();
        ms_bottom_up(a, kmask & a.ctl[MS_ACT0 + (L & 1)], lds.hub, nhub, fm_in, inline_update ? fm_out : nullptr,
                     L + 1, kmask);
        return;
    }
    const uint32_t ntiles = static_cast<uint32_t>((uint64_t(E) + T - 1) / T);
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t tid = threadIdx.x;
        const uint32_t base = t * T;
        const uint64_t tile_end = uint64_t(base) + T;
        const uint32_t end = tile_end < E ? static_cast<uint32_t>(tile_end) : E;
        const uint32_t i0 = a.tile_start[t];
        const uint32_t i1 = (t + 1 < ntiles) ? a.tile_start[t + 1] : F - 1;
        uint32_t c = i1 - i0 + 1;
        if (i1 < i0 || c > T + 1 || i1 >= F) {
            if (tid == 0) atomicOr(&a.ctl[MS_ERR], MS_ERR_TILE);
            c = 0;
        }
        for (uint32_t j = tid; j <= c; j += BLOCK) {
            if (j < c) {
                s.eoff[j] = a.q_eoff[i0 + j];
                s.rb[j] = a.q_rb[i0 + j];
                s.fm[j] = a.q_fm[i0 + j];
            } else {
                s.eoff[j] = (i0 + c < F) ? a.q_eoff[i0 + c] : E;
            }
        }
        __syncthreads();
        if (c) {
            // Edge e belongs to entry o iff eoff[o] <= e < eoff[o + 1];
            // entries and columns failing the range checks are dropped and
            // reported.
            int32_t v[K];
            unsigned long long fm[K];
            bool bad = false;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t e = base + k * BLOCK + tid;
                v[k] = -1;
                fm[k] = 0;
                if (e < end) {
                    const uint32_t o = owner_of(s.eoff, c, e);
                    const uint64_t pos = static_cast<uint64_t>(s.rb[o]) + (e - s.eoff[o]);
                    if (e < s.eoff[o] || e >= s.eoff[o + 1] || pos >= static_cast<uint64_t>(a.m)) {
                        bad = true;
                    } else {
                        fm[k] = s.fm[o];
                        v[k] = a.cols[pos];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (v[k] < -1 || v[k] >= a.n) {
                    v[k] = -1;
                    bad = true;
                }
            }
            if (bad) atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
            unsigned long long sv[K], nv0[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                sv[k] = nv0[k] = ~0ull;
                if (v[k] >= 0) {
                    const SeenNext r = a.sn[v[k]];
                    sv[k] = r.seen;
                    nv0[k] = r.next;
                }
            }
            if (!sparse) {
                // Bits already in next (as far as this load sees: next only
                // gains bits during a level) need no atomic.
                unsigned long long nx[K];
#pragma unroll
                for (int k = 0; k < K; ++k) nx[k] = (v[k] >= 0 && (fm[k] & ~sv[k])) ? nv0[k] : ~0ull;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const unsigned long long mk = fm[k] & ~sv[k] & ~nx[k];
                    if (mk) __hip_atomic_fetch_or(&a.sn[v[k]].next, mk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                bool first[K];
                unsigned long long nx[K];  // as in the dense case; a vertex with bits is listed already
#pragma unroll
                for (int k = 0; k < K; ++k) nx[k] = (v[k] >= 0 && (fm[k] & ~sv[k])) ? nv0[k] : ~0ull;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const unsigned long long mk = fm[k] & ~sv[k] & ~nx[k];
                    first[k] = false;
                    if (mk)
                        first[k] = __hip_atomic_fetch_or(&a.sn[v[k]].next, mk, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT) == 0;
                }
                // Workgroup-aggregated appends: one counter atomic per tile
                // (one per wave and k serialised on the counter when a
                // sparse level touches millions of vertices).  The scan's
                // words reuse the tile's LDS once every wave is past it.
                uint32_t mine = 0;
#pragma unroll
                for (int k = 0; k < K; ++k) mine += first[k] ? 1u : 0u;
                __syncthreads();
                uint32_t* const s_scan = reinterpret_cast<uint32_t*>(lds.hub);
                unsigned long long* const s_pos = lds.hub + BLOCK / 64;
                uint32_t tot;
                const uint32_t excl = block_exclusive_scan<BLOCK>(mine, s_scan, tot);
                if (tid == 0) *s_pos = tot ? atomicAdd(&a.ctl[MS_TCNT0 + npar], static_cast<unsigned long long>(tot)) : 0ull;
                __syncthreads();
                unsigned long long slot = *s_pos + excl;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (first[k]) {
                        if (slot < static_cast<unsigned long long>(a.n)) a.touched[slot] = v[k];
                        else atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
                        ++slot;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Deep, thin batches (road-like graphs: thousands of levels of a few
// thousand vertices per source).  There a level's cost is its chain of
// dependent memory round trips, so the update and the expansion of a level
// run fused in one launch: a thread takes one vertex of the level's touched
// list (loaded beside the list's length), takes its new bits (next & ~seen),
// writes its level bytes, and pushes the new bits along its row -- a probe
// of the target's seen / next, an atomicOr whose old value tells whether the
// target was untouched, and one reservation per wave to append the newly
// touched targets to the next list.  Rows longer than kMsDeepRow go to the
// wave, 64 entries per step.  `next` and the touched list are double-buffered
// by level parity (even levels read sn's next and the touched list, odd ones
// next1 and touched2, each writing the other), so the bits and list of level
// L+1 never meet those being consumed; list lengths are triple-buffered
// (dcnt[L % 3]: level L reads it, writes dcnt[(L+1) % 3] and zeroes
// dcnt[(L+2) % 3], last read by level L-1).  A level whose list is empty ends
// the batch.  No frontier masks or bottom-up state are kept: a fused batch
// runs top-down only.
constexpr int kMsDeepBlock = 256;
constexpr int kMsDeepRow = 8;
constexpr int kMsDeepDepth = 64;  // levels of a batch from which fused levels are learned

__global__ __launch_bounds__(kMsDeepBlock) void ms_deep_kernel(MsArgs a, int L) {
    const uint32_t lane = threadIdx.x & 63u;
    const bool odd = (L & 1) != 0;
    const int32_t* const tin = odd ? a.touched2 : a.touched;
    int32_t* const tout = odd ? a.touched : a.touched2;
    const uint64_t n = static_cast<uint64_t>(a.n);
    const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kMsDeepBlock + threadIdx.x;
    const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kMsDeepBlock;
    const int32_t v_first = gtid < n ? tin[gtid] : -1;  // used only below the list's length
    const unsigned long long stop = a.ctl[MS_DONE] | a.ctl[MS_ERR];
    const unsigned long long F64 = a.ctl[MS_DCNT0 + L % 3];
    if (stop) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl[MS_DCNT0 + (L + 2) % 3] = 0;
        if (F64 == 0) a.ctl[MS_DONE] = 1;
        else atomicMax(&a.ctl[MS_MAXF], F64);
    }
    if (F64 == 0) return;
    if (F64 > n) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
        return;
    }
    const uint32_t F = static_cast<uint32_t>(F64);
    int32_t* const levels = a.p->levels;
    unsigned long long* const cnt_next = &a.ctl[MS_DCNT0 + (L + 1) % 3];
    const unsigned long long lv = static_cast<unsigned long long>(L < line_levels(a) ? L : line_levels(a));
    bool bad = false, anynew = false;
    // next of level L+1 (the words this level ORs into)
    auto next_out = [&](int64_t w) -> unsigned long long* { return odd ? &a.sn[w].next : &a.next1[w]; };
    // Reserves c list slots for each lane (wave-uniform call, one atomicAdd
    // per wave); the lane's first slot, or ~0u if the list would overflow.
    auto reserve = [&](uint32_t c) -> uint32_t {
        const uint32_t incl = wave_inclusive_scan(c);
        const uint32_t tot = __shfl(incl, 63, 64);
        if (tot == 0) return 0u;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(cnt_next, static_cast<unsigned long long>(tot));
        base = __shfl(base, 0, 64);
        if (base + tot > n) {
            bad = true;
            return ~0u;
        }
        return static_cast<uint32_t>(base) + incl - c;
    };
    // The same for the whole workgroup (every thread calls it): one block
    // scan and one atomicAdd.  A per-wave reservation on the one counter
    // serialises when a level touches many vertices (64-source batches).
    __shared__ uint32_t s_cnt[kMsDeepBlock / 64];
    __shared__ unsigned long long s_base;
    auto reserve_block = [&](uint32_t c) -> uint32_t {
        uint32_t tot;
        const uint32_t excl = block_exclusive_scan<kMsDeepBlock>(c, s_cnt, tot);
        if (threadIdx.x == 0) s_base = tot ? atomicAdd(cnt_next, static_cast<unsigned long long>(tot)) : 0ull;
        __syncthreads();
        const unsigned long long base = s_base;
        __syncthreads();  // s_base is written again by the next call
        if (tot == 0) return 0u;
        if (base + tot > n) {
            bad = true;
            return ~0u;
        }
        return static_cast<uint32_t>(base) + excl;
    };
    // Workgroup-uniform steps: the workgroup's threads hold consecutive
    // entries [base, base + kMsDeepBlock).
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kMsDeepBlock; base < F; base += gstride) {
        const uint64_t i = base + threadIdx.x;
        const bool has = i < F;
        int64_t v = -1;
        if (has) {
            const int32_t t = i == gtid ? v_first : tin[i];
            if (static_cast<uint64_t>(static_cast<uint32_t>(t)) >= n) bad = true;
            else v = t;
        }
        // The vertex's state and row bounds.
        unsigned long long sv = 0, nx = 0;
        int64_t rb = 0, re = 0;
        if (v >= 0) {
            const ulonglong2 st = *reinterpret_cast<const ulonglong2*>(a.sn + v);
            sv = st.x;
            nx = odd ? a.next1[v] : st.y;
            rb = a.offsets[v];
            re = a.offsets[v + 1];
        }
        const unsigned long long nv = nx & ~sv;
        const bool first = nv && !sv;
        const uint32_t deg = static_cast<uint32_t>(re - rb);
        const bool short_row = nv && deg <= kMsDeepRow;
        // Its line (when an earlier level wrote it) and, for a short row, its
        // columns and their targets' state.
        ulonglong2 line[4];
        const ulonglong2* src = line_of(a, v >= 0 ? v : 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            line[j] = (nv && !first && j < line_pairs(a)) ? src[j] : make_ulonglong2(~0ull, ~0ull);
        int32_t u[kMsDeepRow];
#pragma unroll
        for (int j = 0; j < kMsDeepRow; ++j) {
            u[j] = (short_row && static_cast<uint32_t>(j) < deg) ? a.cols[rb + j] : -1;
            if (u[j] >= 0 && static_cast<uint64_t>(u[j]) >= n) {
                bad = true;
                u[j] = -1;
            }
        }
        unsigned long long mk[kMsDeepRow];
#pragma unroll
        for (int j = 0; j < kMsDeepRow; ++j) {
            mk[j] = 0;
            if (u[j] >= 0) {
                const ulonglong2 t = *reinterpret_cast<const ulonglong2*>(a.sn + u[j]);
                mk[j] = nv & ~t.x & ~(odd ? t.y : a.next1[u[j]]);
            }
        }
        // The vertex's own state: only words no other thread of this level
        // writes (its seen and its next of level L).
        if (v >= 0 && nx) {
            if (odd) {
                a.sn[v].seen = sv | nv;
                a.next1[v] = 0;
            } else {
                *reinterpret_cast<ulonglong2*>(a.sn + v) = make_ulonglong2(sv | nv, 0ull);
            }
        }
        if (first) atomicOr(&a.reached[v >> 6], 1ull << (v & 63));
        if (L >= line_levels(a)) {
            // every lane of the wave takes part in the shuffles of write_levels
            if (__ballot(nv != 0)) write_levels(levels, a.n, nv ? a.iperm[v] : 0, nv, L);
        }
        if (nv) {
            anynew = true;
            ulonglong2* dst = const_cast<ulonglong2*>(line_of(a, v));
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < line_pairs(a)) dst[j] = merge_line_pair(a, line[j], nv, j, lv);
        }
        // Pushes along a short row; targets it turns from untouched to
        // touched join the next list.
        uint32_t claim = 0;
#pragma unroll
        for (int j = 0; j < kMsDeepRow; ++j)
            if (mk[j] && __hip_atomic_fetch_or(next_out(u[j]), mk[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                claim |= 1u << j;
        const uint32_t slot0 = reserve_block(static_cast<uint32_t>(__popc(claim)));
        if (claim && slot0 != ~0u) {
            uint32_t r = 0;
#pragma unroll
            for (int j = 0; j < kMsDeepRow; ++j)
                if ((claim >> j) & 1u) tout[slot0 + r++] = u[j];
        }
        // Long rows, one at a time by the whole wave.
        uint64_t lm = wave_ballot(nv && deg > kMsDeepRow);
        while (lm) {
            const int hl = __ffsll(static_cast<long long>(lm)) - 1;
            lm &= lm - 1;
            const int64_t hrb = __shfl(rb, hl, 64), hre = __shfl(re, hl, 64);
            const unsigned long long hnv = __shfl(nv, hl, 64);
            for (int64_t b = hrb; b < hre; b += 64) {
                int32_t w = b + lane < hre ? a.cols[b + lane] : -1;
                if (w >= 0 && static_cast<uint64_t>(w) >= n) {
                    bad = true;
                    w = -1;
                }
                unsigned long long m = 0;
                if (w >= 0) {
                    const ulonglong2 t = *reinterpret_cast<const ulonglong2*>(a.sn + w);
                    m = hnv & ~t.x & ~(odd ? t.y : a.next1[w]);
                }
                const bool c = m && __hip_atomic_fetch_or(next_out(w), m, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT) == 0;
                const uint32_t slot = reserve(c ? 1u : 0u);
                if (c && slot != ~0u) tout[slot] = w;
            }
        }
    }
    // Every writer of this level stores the same L and levels run in order,
    // so the final value is the deepest level reached.
    if (wave_ballot(anynew) && lane == 0) a.ctl[MS_MAXLEV] = static_cast<unsigned long long>(L);
    if (bad) atomicOr(&a.ctl[MS_ERR], MS_ERR_RANGE);
}

using MsExpandFn = void (*)(MsArgs, int);
struct MsKernelSet {
    int block, k;
    MsExpandFn expand;
};
#define GG_MSKSET(B, K) {B, K, ms_expand_kernel<B, K>}
const MsKernelSet kMsKernels[] = {GG_MSKSET(256, 1), GG_MSKSET(256, 2), GG_MSKSET(256, 4), GG_MSKSET(256, 8),
                                  GG_MSKSET(512, 1), GG_MSKSET(512, 2), GG_MSKSET(512, 4), GG_MSKSET(1024, 1),
                                  GG_MSKSET(1024, 2), GG_MSKSET(128, 4), GG_MSKSET(128, 8), GG_MSKSET(64, 8),
                                  GG_MSKSET(64, 16)};
#undef GG_MSKSET

const MsKernelSet* find_ms_kernels(int block, int k) {
    for (const auto& ks : kMsKernels)
        if (ks.block == block && ks.k == k) return &ks;
    return nullptr;
}

// Per-source statistics from the final seen masks: reached vertices and the
// sum of their degrees, reduced per wave then per workgroup into
// partial[blockIdx][k][2]; the host adds the rows.
__global__ __launch_bounds__(256) void ms_stats_kernel(MsArgs a, unsigned long long* partial) {
    __shared__ unsigned long long s_acc[64][2];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) s_acc[i][0] = s_acc[i][1] = 0;
    __syncthreads();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < a.n; base += stride) {
        const int64_t v = base + threadIdx.x;
        unsigned long long sv = 0;
        unsigned long long d = 0;
        if (v < a.n) {
            sv = a.sn[v].seen;
            d = static_cast<unsigned long long>(a.offsets[v + 1] - a.offsets[v]);
        }
        unsigned long long any = sv;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) any |= __shfl_xor(any, o, 64);
        while (any) {
            const int s = __ffsll(static_cast<long long>(any)) - 1;
            any &= any - 1;
            const bool has = (sv >> s) & 1ull;
            const unsigned long long cntw = __popcll(__ballot(has));
            const unsigned long long degw = wave_sum(has ? d : 0ull);
            if ((threadIdx.x & 63) == 0) {
                atomicAdd(&s_acc[s][0], cntw);
                atomicAdd(&s_acc[s][1], degw);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.p->k; i += blockDim.x) {
        partial[(int64_t(blockIdx.x) * 64 + i) * 2 + 0] = s_acc[i][0];
        partial[(int64_t(blockIdx.x) * 64 + i) * 2 + 1] = s_acc[i][1];
    }
}

constexpr int kStatsBlocks = 1024;

}  // namespace

bool ms_config_supported(int block, int k) { return find_ms_kernels(block, k) != nullptr; }

namespace {

// First new id of each degree key (key 0 = the highest degree bucket).
struct KeyBounds {
    int64_t first[kDegreeBuckets + 1];
};

__device__ __forceinline__ int degree_key(const KeyBounds& kb, int32_t u) {
    int lo = 0, hi = kDegreeBuckets - 1;  // last key whose first id is <= u
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (kb.first[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Row i of the degree-ordered copy = row iperm[i] of the graph, columns
// mapped through perm and ordered by the neighbours' degree key (highest
// degrees first, file order inside a key): a bottom-up scan meets the hubs,
// the vertices most likely to carry frontier bits, first and so stops
// earlier.  One wave per row, a stable counting sort over the keys: one
// pass counts (LDS atomics), a wave scan turns the counts into bases, a
// second pass places 64 entries at a time, key by key, with ballots.
constexpr int kRelabelBlock = 256;
__global__ __launch_bounds__(kRelabelBlock) void ms_relabel_kernel(
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ cols, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ iperm, const int64_t* __restrict__ new_offsets, int32_t* __restrict__ new_cols,
    int64_t n, KeyBounds kb) {
    __shared__ uint32_t s_base[kRelabelBlock / 64][kDegreeBuckets];
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* base = s_base[threadIdx.x >> 6];
    const int64_t nwaves = int64_t(gridDim.x) * (blockDim.x / 64);
    for (int64_t i = int64_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); i < n; i += nwaves) {
        const int64_t o = iperm[i];
        const int64_t rb = offsets[o], re = offsets[o + 1], nb = new_offsets[i];
        if (re == rb) continue;
        if (lane < kDegreeBuckets) base[lane] = 0;
        wave_lds_sync();
        for (int64_t c0 = rb; c0 < re; c0 += 64) {
            const int64_t j = c0 + lane;
            if (j < re) atomicAdd(&base[degree_key(kb, perm[cols[j]])], 1u);
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
            const int32_t u = valid ? perm[cols[j]] : 0;
            const int key = valid ? degree_key(kb, u) : -1;
            uint64_t todo = wave_ballot(valid);
            while (todo) {
                const int leader = __ffsll(static_cast<long long>(todo)) - 1;
                const int kk = __shfl(key, leader, 64);
                const uint64_t mk = wave_ballot(key == kk);
                const uint32_t b = base[kk];
                if (key == kk) new_cols[nb + b + rank_below(mk)] = u;
                wave_lds_sync();
                if (static_cast<int>(lane) == leader) base[kk] = b + static_cast<uint32_t>(__popcll(mk));
                wave_lds_sync();
                todo &= ~mk;
            }
        }
    }
}

}  // namespace

void build_ms_graph(DeviceGraph& dg, const HostGraph& g) {
    const int64_t n = g.n, m = g.m();
    // Counting sort of the vertices by descending degree bucket, ids
    // ascending inside a bucket.
    std::vector<int64_t> start(kDegreeBuckets + 1, 0);
    std::vector<uint8_t> bucket(static_cast<size_t>(n));
    for (int64_t v = 0; v < n; ++v) {
        bucket[v] = static_cast<uint8_t>(degree_bucket(g.offsets[v + 1] - g.offsets[v]));
        ++start[kDegreeBuckets - 1 - bucket[v] + 1];
    }
    for (int b = 0; b < kDegreeBuckets; ++b) start[b + 1] += start[b];
    KeyBounds kb;
    for (int b = 0; b <= kDegreeBuckets; ++b) kb.first[b] = start[b];
    std::vector<int32_t> perm(static_cast<size_t>(n)), iperm(static_cast<size_t>(n));
    for (int64_t v = 0; v < n; ++v) {
        const int64_t pos = start[kDegreeBuckets - 1 - bucket[v]]++;
        iperm[pos] = static_cast<int32_t>(v);
        perm[v] = static_cast<int32_t>(pos);
    }
    // Heavy rows lead (descending degree buckets); order them by exact
    // descending degree, so that the rows with a tile in round r (tile r of
    // every row that long) are a prefix.  Tiles are numbered round by round:
    // the first tiles of all heavy rows come first, and a row covered by
    // them skips its later tiles.
    auto degree_of = [&](int32_t v) { return g.offsets[v + 1] - g.offsets[v]; };
    int64_t nh = 0;
    while (nh < n && degree_of(iperm[nh]) >= kMsHeavyRow) ++nh;
    std::stable_sort(iperm.begin(), iperm.begin() + nh,
                     [&](int32_t x, int32_t y) { return degree_of(x) > degree_of(y); });
    for (int64_t i = 0; i < nh; ++i) perm[iperm[i]] = static_cast<int32_t>(i);
    std::vector<int64_t> noff(static_cast<size_t>(n) + 1, 0);
    for (int64_t i = 0; i < n; ++i) noff[i + 1] = noff[i] + degree_of(iperm[i]);
    const int32_t rounds = nh ? static_cast<int32_t>((degree_of(iperm[0]) + kMsHeavyRow - 1) / kMsHeavyRow) : 0;
    std::vector<uint32_t> tiles(static_cast<size_t>(rounds) + 1, 0);
    for (int32_t r = 0, rows = static_cast<int32_t>(nh); r < rounds; ++r) {
        while (rows > 0 && degree_of(iperm[rows - 1]) <= int64_t(r) * kMsHeavyRow) --rows;
        tiles[r + 1] = tiles[r] + static_cast<uint32_t>(rows);
    }
    dg.ms_heavy = nh;
    dg.ms_heavy_rounds = rounds;
    // On a symmetric graph a vertex without edges has no in-edges either: it
    // is reached only as a source, at level 0, and the scans skip it.  On a
    // directed graph a vertex without out-edges can be reached at any level.
    int64_t nz = n;
    if (dg.symmetric)
        while (nz > 0 && degree_of(iperm[nz - 1]) == 0) --nz;
    dg.ms_nonzero = nz;
    GG_HIP(hipMalloc(&dg.ms_heavy_tiles, sizeof(uint32_t) * (rounds + 1)));
    dg.graph_bytes += sizeof(uint32_t) * (rounds + 1);
    GG_HIP(hipMemcpy(dg.ms_heavy_tiles, tiles.data(), sizeof(uint32_t) * (rounds + 1), hipMemcpyHostToDevice));
    GG_HIP(hipMalloc(&dg.ms_perm, sizeof(int32_t) * n));
    GG_HIP(hipMalloc(&dg.ms_iperm, sizeof(int32_t) * n));
    GG_HIP(hipMalloc(&dg.ms_offsets, sizeof(int64_t) * (n + 1)));
    GG_HIP(hipMalloc(&dg.ms_cols, sizeof(int32_t) * std::max<int64_t>(m, 1)));
    dg.graph_bytes += sizeof(int32_t) * 2 * n + sizeof(int64_t) * (n + 1) + sizeof(int32_t) * std::max<int64_t>(m, 1);
    GG_HIP(hipMemcpy(dg.ms_perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    GG_HIP(hipMemcpy(dg.ms_iperm, iperm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    GG_HIP(hipMemcpy(dg.ms_offsets, noff.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    if (m) {
        const uint32_t grid = static_cast<uint32_t>(std::min<int64_t>((n + 3) / 4, int64_t(dg.num_cus) * 8));
        hipLaunchKernelGGL(ms_relabel_kernel, dim3(grid), dim3(kRelabelBlock), 0, nullptr, dg.offsets, dg.cols,
                           dg.ms_perm, dg.ms_iperm, dg.ms_offsets, dg.ms_cols, n, kb);
        GG_HIP(hipGetLastError());
        GG_HIP(hipStreamSynchronize(nullptr));
    }
}

void Runner::free_ms_workspace() {
    if (ms.graph_exec) (void)hipGraphExecDestroy(ms.graph_exec);
    void* ptrs[] = {ms.sn,      ms.fmask,      ms.lv8,     ms.reached, ms.q_rb,   ms.q_eoff, ms.q_v,
                    ms.q_fm,    ms.tile_start, ms.touched, ms.ctl,     ms.params, ms.levels, ms.partial,
                    ms.touched2, ms.next1};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (ms.h_params) (void)hipHostFree(ms.h_params);
    if (ms.h_ctl) (void)hipHostFree(ms.h_ctl);
    ms = MsWorkspace{};
}

void Runner::ensure_ms_workspace() {
    // Sized for the smallest tile of any supported shape: a change of shape
    // only rebuilds the captured sequence (its key includes the shape).
    int T = kMsKernels[0].block * kMsKernels[0].k;
    for (const auto& ks : kMsKernels) T = std::min(T, ks.block * ks.k);
    const int64_t tile_cap = (m + T - 1) / T + 2;
    if (ms.ctl) return;
    const int64_t nchunks = (n + kUcChunk - 1) / kUcChunk;
    size_t bytes = 0;
    auto alloc = [&](auto** p, size_t sz) {
        GG_HIP(hipMalloc(reinterpret_cast<void**>(p), sz));
        bytes += sz;
    };
    try {
        alloc(&ms.sn, sizeof(SeenNext) * n);
        alloc(&ms.fmask, sizeof(unsigned long long) * n);
        alloc(&ms.lv8, sizeof(unsigned long long) * 8 * n);
        alloc(&ms.reached, sizeof(unsigned long long) * ((n + 63) / 64));
        alloc(&ms.q_rb, sizeof(int64_t) * n);
        alloc(&ms.q_eoff, sizeof(uint32_t) * n);
        alloc(&ms.q_v, sizeof(int32_t) * n);
        alloc(&ms.q_fm, sizeof(unsigned long long) * n);
        alloc(&ms.tile_start, sizeof(uint32_t) * tile_cap);
        alloc(&ms.touched, sizeof(int32_t) * n);
        alloc(&ms.touched2, sizeof(int32_t) * n);
        alloc(&ms.next1, sizeof(unsigned long long) * n);
        // The second mask copy shares next1's memory: only fused deep batches
        // use next1 and only the others fmask2; the clear zeroes next1 at the
        // start of every batch, and fmask2 is written whole before it is read.
        ms.fmask2 = ms.next1;
        alloc(&ms.ctl, sizeof(unsigned long long) * kCtlWords);
        alloc(&ms.params, sizeof(MsParams));
        alloc(&ms.partial, sizeof(unsigned long long) * kStatsBlocks * 64 * 2);
        // Coherent host memory: the clear kernel reads the parameters from
        // it, and the finish or snapshot kernel stores the control block.
        GG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ms.h_params), sizeof(MsParams),
                             hipHostMallocCoherent | hipHostMallocMapped));
        GG_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&ms.h_params_dev), ms.h_params, 0));
        GG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ms.h_ctl), sizeof(unsigned long long) * 2 * kCtlWords,
                             hipHostMallocCoherent | hipHostMallocMapped));
        GG_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&ms.h_ctl_dev), ms.h_ctl, 0));
    } catch (...) {
        free_ms_workspace();
        throw;
    }
    ms.tile_cap = tile_cap;
    ms.tile_edges = T;
    ms.nchunks = nchunks;
    ms.bytes = bytes;
}

void Runner::prepare_batch() {
    DeviceGuard guard(device);
    ensure_ms_workspace();
}

// What ensure_ms_workspace() allocates: per vertex 13 words (sn 2, fmask,
// lv8 8, q_fm, next1 = fmask2), q_rb, and four 4-byte arrays (q_eoff, q_v,
// touched, touched2); the reached bitmap; tile indices; the statistics.
size_t batch_bytes_estimate(int64_t n_vertices, int64_t m_entries, int k, bool host_levels) {
    size_t b = sizeof(unsigned long long) * n_vertices * 13 + (sizeof(int64_t) + 16) * n_vertices +
               sizeof(unsigned long long) * ((n_vertices + 63) / 64) +
               sizeof(uint32_t) * (m_entries / 256 + 2) + 1024 * 64 * 16 + 4096;
    if (host_levels) b += sizeof(int32_t) * n_vertices * k;
    return b;
}

void Runner::bfs_batch(const int32_t* h_sources, int k, int32_t* d_levels, gb_batch_stats* st) {
    static_assert(kCtlWords >= MS_WORDS, "control block size");
    if (k < 1 || k > 64) fail(GB_EINVAL, "a batch holds 1 to 64 sources");
    if (!h_sources || !d_levels) fail(GB_EINVAL, "null sources or levels");
    for (int i = 0; i < k; ++i)
        if (h_sources[i] < 0 || h_sources[i] >= n)
            fail(GB_ERANGE, "source " + std::to_string(h_sources[i]) + " outside [0, " + std::to_string(n) + ")");
    const MsKernelSet* ks = find_ms_kernels(g.ms_cfg.block, g.ms_cfg.edges_per_thread);
    if (!ks) fail(GB_EINVAL, "unsupported batch block / edges_per_thread combination");
    DeviceGuard guard(device);
    ensure_ms_workspace();

    const int64_t sparse_cfg = g.ms_cfg.sparse_edges > 0 ? g.ms_cfg.sparse_edges : g.default_ms_sparse_edges();
    MsArgs a;
    a.offsets = g.ms_offsets;
    a.cols = g.ms_cols;
    a.perm = g.ms_perm;
    a.iperm = g.ms_iperm;
    a.heavy_tiles = g.ms_heavy_tiles;
    a.heavy = g.ms_heavy;
    a.heavy_rounds = g.ms_heavy_rounds;
    a.nz = g.ms_nonzero;
    a.nz_chunks = static_cast<uint32_t>((g.ms_nonzero + kUcChunk - 1) / kUcChunk);
    a.sn = reinterpret_cast<SeenNext*>(ms.sn);
    a.fmask = ms.fmask;
    a.fmask2 = ms.fmask2;
    a.lv8 = ms.lv8;
    a.reached = ms.reached;
    a.p = ms.params;
    a.hp = ms.h_params_dev;
    a.q_rb = ms.q_rb;
    a.q_eoff = ms.q_eoff;
    a.q_v = ms.q_v;
    a.q_fm = ms.q_fm;
    a.tile_start = ms.tile_start;
    a.touched = ms.touched;
    a.touched2 = ms.touched2;
    a.next1 = ms.next1;
    a.ctl = ms.ctl;
    a.n = n;
    a.m = m;
    a.nchunks = static_cast<uint32_t>(ms.nchunks);
    a.tile_edges = static_cast<uint32_t>(g.ms_cfg.block * g.ms_cfg.edges_per_thread);
    a.sparse_edges = static_cast<uint32_t>(std::min<int64_t>(sparse_cfg, 0xffffffffll));
    a.bottom_up = (g.symmetric && g.ms_cfg.direction_opt) ? 1u : 0u;
    a.alpha = static_cast<uint32_t>(g.ms_cfg.alpha);
    a.low_degree = static_cast<uint32_t>(g.ms_cfg.low_degree);
    a.beta = static_cast<uint32_t>(g.ms_cfg.beta);
    // Copying the hub masks costs every workgroup 10 KB per wave of reads per
    // bottom-up level: worth it on large graphs only.
    a.hub_lds = n >= kHubLdsMinVertices ? 1u : 0u;

    int occ = 0;
    GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(ks->expand),
                                                       g.ms_cfg.block, 0));
    if (occ < 1) fail(GB_EDEVICE, "batch kernel cannot be resident on a CU");
    const int waves_per_block = (g.ms_cfg.block + 63) / 64;
    const int per_cu = std::max(1, std::min(occ, 16 / waves_per_block));
    const uint32_t grid = static_cast<uint32_t>(per_cu * g.num_cus);
    int uocc = 0;
    GG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&uocc, reinterpret_cast<const void*>(ms_update_compact_kernel),
                                                       kUcBlock, 0));
    // At least 2 workgroups per CU whatever the occupancy query reports (it
    // grants 1 for this kernel's registers): chunks are independent, and the
    // extra workgroups cost nothing where they cannot be resident (R-MAT s24
    // 64-source batch 7.51 -> 7.47 ms, s20 0.955 -> 0.944 ms).
    const uint32_t ugrid = static_cast<uint32_t>(
        std::max<int64_t>(1, std::min<int64_t>(ms.nchunks, int64_t(g.num_cus) * std::max(uocc, 2))));
    // finish: up to 64 workgroups per CU (R-MAT s24, 64-source batch: 8.08 ms
    // against 8.24 ms with 8 per CU, 8.15 with 32, 8.10 with one 64-vertex
    // group per wave).
    const uint32_t fgrid = static_cast<uint32_t>(std::min<int64_t>((n + 255) / 256, int64_t(g.num_cus) * 64));
    const uint32_t cgrid = static_cast<uint32_t>(std::min<int64_t>((n * 16 + 255) / 256 + 1, int64_t(g.num_cus) * 8));

    std::memset(ms.h_params, 0, sizeof(MsParams));
    std::memcpy(ms.h_params->sources, h_sources, sizeof(int32_t) * k);
    ms.h_params->k = k;
    ms.h_params->levels = d_levels;
    // A fresh workspace, or one a failed batch left, is cleared whole.
    ms.h_params->clear_all = ms.tail_clean ? 0 : 1;
    std::memcpy(ms.h_params->prev_sources, ms.prev_sources, sizeof(ms.prev_sources));
    ms.h_params->prev_k = ms.tail_clean ? ms.prev_k : 0;
    // The window holds for the direction thresholds it was learned with.
    const bool window = ms.bu_alpha == g.ms_cfg.alpha && ms.bu_beta == g.ms_cfg.beta;
    ms.h_params->bu_lo = window ? ms.bu_lo : 1;
    ms.h_params->bu_hi = window ? ms.bu_hi : 0;
    ms.tail_clean = false;
    auto launch_prologue = [&]() {
        hipLaunchKernelGGL(ms_clear_kernel, dim3(cgrid), dim3(256), 0, stream, a);
        hipLaunchKernelGGL(ms_seed_kernel, dim3(1), dim3(64), 0, stream, a);
    };
    // Fused levels (ms_deep_kernel) for batches on graphs whose last batch
    // was deep (>= kMsDeepDepth levels) with thin, top-down frontiers only.
    const bool deep = g.ms_cfg.deep == GB_DEEP_ON || (g.ms_cfg.deep == GB_DEEP_AUTO && ms.deep_learned);
    const uint32_t dgrid = static_cast<uint32_t>(
        std::max<int64_t>(1, std::min<int64_t>((n + kMsDeepBlock - 1) / kMsDeepBlock, int64_t(g.num_cus) * 2)));
    // Nibble level lines when the last batch of this runner was shallow
    // enough for them (a deeper batch stays exact: its deeper levels are
    // stored in the output rows directly, and the next batch uses bytes).
    a.nib = (!deep && ms.depth_hint > 0 && ms.depth_hint <= kNibLevels) ? 1u : 0u;
    int L = 0;
    auto launch_level = [&]() {
        if (deep) {
            hipLaunchKernelGGL(ms_deep_kernel, dim3(dgrid), dim3(kMsDeepBlock), 0, stream, a, L);
        } else {
            hipLaunchKernelGGL(ms_update_compact_kernel, dim3(ugrid), dim3(kUcBlock), 0, stream, a, L);
            hipLaunchKernelGGL(ks->expand, dim3(grid), dim3(g.ms_cfg.block), 0, stream, a, L);
        }
        ++L;
    };

    const int depth_hint = std::min(std::max(ms.depth_hint, 4), kMaxCapturedLevels);
    // The finish runs right after the first depth_hint levels when they have
    // ended the search (the kernel checks the done flag), and the control
    // block is snapshotted behind it: a batch that fits the hint costs the
    // host one wait.
    // The finish also stores the control block to h_ctl slot `slot`.
    auto launch_finish = [&](int only_if_done, int slot) {
        unsigned long long* snap = ms.h_ctl_dev + slot * kCtlWords;
        if (a.nib)
            hipLaunchKernelGGL(ms_finish_kernel<true>, dim3(fgrid), dim3(256), 0, stream, a, only_if_done, snap);
        else
            hipLaunchKernelGGL(ms_finish_kernel<false>, dim3(fgrid), dim3(256), 0, stream, a, only_if_done, snap);
    };
    auto snapshot_to = [&](int slot) {
        hipLaunchKernelGGL(ms_snapshot_kernel, dim3(1), dim3(64), 0, stream,
                           static_cast<const unsigned long long*>(ms.ctl), ms.h_ctl_dev + slot * kCtlWords);
    };
    auto first_part = [&]() {
        launch_prologue();
        for (int i = 0; i < depth_hint; ++i) launch_level();
        launch_finish(1, 0);
    };
    GG_HIP(hipEventRecord(ev_start, stream));
    if (g.ms_cfg.use_graph) {
        // One captured sequence: parameter upload, clears, depth_hint levels,
        // the conditional finish and the snapshot.  Everything batch-specific
        // is read from device memory, so the graph is rebuilt only when the
        // kernel shape or the depth hint changes.
        const bool reuse = ms.graph_exec && ms.graph_depth == depth_hint && ms.graph_block == g.ms_cfg.block &&
                           ms.graph_kpt == g.ms_cfg.edges_per_thread && ms.graph_sparse == a.sparse_edges &&
                           ms.graph_bu == int(a.bottom_up) && ms.graph_alpha == int(a.alpha) &&
                           ms.graph_low == int(a.low_degree) && ms.graph_deep == int(deep) &&
                           ms.graph_beta == int(a.beta) && ms.graph_nib == int(a.nib);
        if (!reuse) {
            if (ms.graph_exec) (void)hipGraphExecDestroy(ms.graph_exec);
            ms.graph_exec = nullptr;
            hipGraph_t captured = nullptr;
            GG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                first_part();
            } catch (...) {
                hipGraph_t dummy = nullptr;
                (void)hipStreamEndCapture(stream, &dummy);
                if (dummy) (void)hipGraphDestroy(dummy);
                throw;
            }
            GG_HIP(hipStreamEndCapture(stream, &captured));
            hipError_t e = hipGraphInstantiate(&ms.graph_exec, captured, nullptr, nullptr, 0);
            (void)hipGraphDestroy(captured);
            GG_HIP(e);
            ms.graph_depth = depth_hint;
            ms.graph_block = g.ms_cfg.block;
            ms.graph_kpt = g.ms_cfg.edges_per_thread;
            ms.graph_sparse = a.sparse_edges;
            ms.graph_bu = int(a.bottom_up);
            ms.graph_alpha = int(a.alpha);
            ms.graph_low = int(a.low_degree);
            ms.graph_deep = int(deep);
            ms.graph_beta = int(a.beta);
            ms.graph_nib = int(a.nib);
            GG_HIP(hipEventRecord(ev_start, stream));
        }
        GG_HIP(hipGraphLaunch(ms.graph_exec, stream));
    } else {
        first_part();
    }
    L = depth_hint;
    GG_HIP(hipGetLastError());
    GG_HIP(hipEventRecord(ev_stop, stream));
    GG_HIP(hipEventRecord(ev_batch[0], stream));
    auto done = [&](int slot) {
        GG_HIP(hipEventSynchronize(ev_batch[slot]));
        const unsigned long long* c = ms.h_ctl + slot * kCtlWords;
        return c[MS_DONE] != 0 || c[MS_ERR] != 0;
    };
    int last = 0;  // h_ctl slot of the final control block
    if (!done(0)) {
        // Deeper than the hint: batches of levels of doubling size (at most
        // kMaxLevelBatch) are pipelined -- the done flag of batch b-1 is
        // checked while batch b runs; levels past the end are no-ops.
        auto snapshot = [&](int slot) {
            snapshot_to(slot);
            GG_HIP(hipEventRecord(ev_batch[slot], stream));
        };
        int batch = 8;
        for (int i = 0; i < batch; ++i) launch_level();
        GG_HIP(hipGetLastError());
        snapshot(1);
        for (int b = 2;; ++b) {
            if (L > n + 2 * kMaxLevelBatch) fail(GB_EDEVICE, "internal error: more BFS levels than vertices");
            batch = std::min(batch * 2, kMaxLevelBatch);
            for (int i = 0; i < batch; ++i) launch_level();
            GG_HIP(hipGetLastError());
            snapshot(b & 1);
            if (done((b - 1) & 1)) {
                last = b & 1;
                break;
            }
        }
        // All levels have run (later ones were no-ops): level rows from the
        // byte lines, then the final control block (behind the snapshot
        // still in flight to the same slot).
        launch_finish(0, last);
        GG_HIP(hipGetLastError());
        GG_HIP(hipEventRecord(ev_stop, stream));
        GG_HIP(hipEventRecord(ev_batch[last], stream));
        GG_HIP(hipEventSynchronize(ev_batch[last]));
    }
    const unsigned long long* fin = ms.h_ctl + last * kCtlWords;
    if (fin[MS_ERR] & MS_ERR_TILE) fail(GB_EDEVICE, "internal error: inconsistent frontier tile index");
    if (fin[MS_ERR] & MS_ERR_RANGE)
        fail(GB_EDEVICE, "internal error: frontier entry or column index out of range");
    if (!fin[MS_DONE]) fail(GB_EDEVICE, "internal error: batch did not reach an empty frontier");
    std::memcpy(ms.prev_sources, h_sources, sizeof(int32_t) * k);
    ms.prev_k = k;
    ms.tail_clean = true;
    ms.last_depth = static_cast<int>(fin[MS_MAXLEV]) + 1;
    // Fused levels next when this batch was deep and thin: no bottom-up or
    // dense level (or, fused, no list longer than a sparse level's edges / 2).
    ms.last_deep = deep;
    ms.deep_learned = deep ? fin[MS_MAXF] * 2 <= static_cast<unsigned long long>(a.sparse_edges)
                           : ms.last_depth >= kMsDeepDepth && fin[MS_BULEVELS] == 0 &&
                                 fin[MS_MAXE] <= static_cast<unsigned long long>(a.sparse_edges);
    ms.last_bu_levels = static_cast<int>(fin[MS_BULEVELS]);
    const bool any_bu = fin[MS_BUFIRST] != ~0ull && fin[MS_BUFIRST] <= fin[MS_BULAST];
    ms.bu_lo = any_bu ? static_cast<int>(fin[MS_BUFIRST]) : 1;
    ms.bu_hi = any_bu ? static_cast<int>(fin[MS_BULAST]) : 0;
    ms.bu_alpha = g.ms_cfg.alpha;
    ms.bu_beta = g.ms_cfg.beta;
    // Levels queued for the next batch: this depth + 1 (an empty level ends a
    // batch), raised at once and lowered only after 16 shallower batches in a
    // row, so that batches of slightly different depths replay one capture.
    const int want = ms.last_depth + 1;
    if (want >= ms.depth_hint) {
        ms.depth_hint = want;
        ms.short_streak = 0;
    } else if (++ms.short_streak >= 16) {
        ms.depth_hint = want;
        ms.short_streak = 0;
    }
    float ms_elapsed = 0.f;
    GG_HIP(hipEventElapsedTime(&ms_elapsed, ev_start, ev_stop));
    last_ms = ms_elapsed;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->ms = ms_elapsed;
        st->num_levels = ms.last_depth;
        st->k = k;
        hipLaunchKernelGGL(ms_stats_kernel, dim3(kStatsBlocks), dim3(256), 0, stream, a, ms.partial);
        GG_HIP(hipGetLastError());
        std::vector<unsigned long long> part(size_t(kStatsBlocks) * 64 * 2);
        GG_HIP(hipMemcpyAsync(part.data(), ms.partial, sizeof(unsigned long long) * part.size(),
                              hipMemcpyDeviceToHost, stream));
        GG_HIP(hipStreamSynchronize(stream));
        for (int b = 0; b < kStatsBlocks; ++b)
            for (int i = 0; i < k; ++i) {
                st->n_reached[i] += static_cast<int64_t>(part[(size_t(b) * 64 + i) * 2 + 0]);
                st->m_reached[i] += static_cast<int64_t>(part[(size_t(b) * 64 + i) * 2 + 1]);
            }
    }
}

void Runner::bfs_batch_to_host(const int32_t* h_sources, int k, int32_t* h_levels, gb_batch_stats* st) {
    if (!h_levels) fail(GB_EINVAL, "levels buffer is null");
    if (k < 1 || k > 64) fail(GB_EINVAL, "a batch holds 1 to 64 sources");
    DeviceGuard guard(device);
    const size_t need = sizeof(int32_t) * size_t(n) * size_t(k);
    if (ms.levels_bytes < need) {
        if (ms.levels) (void)hipFree(ms.levels);
        ms.levels = nullptr;
        ms.levels_bytes = 0;
        GG_HIP(hipMalloc(&ms.levels, need));
        ms.levels_bytes = need;
    }
    bfs_batch(h_sources, k, ms.levels, st);
    GG_HIP(hipMemcpyAsync(h_levels, ms.levels, need, hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
}

}  // namespace gg
