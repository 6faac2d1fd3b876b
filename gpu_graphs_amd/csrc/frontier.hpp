// Device helpers shared by the single-source and multi-source BFS kernels:
// the tile index that maps an edge offset of a level to the frontier entry
// owning it.
//
// Frontier queues are built in chunks (a workgroup's worth of entries) that
// each reserve their queue slots and edge range with ONE 64-bit atomicAdd of
// (count << 32 | edges) on the level's counter.  Both halves advance in the
// same atomic, so slot order and edge order agree and the queue's edge
// prefixes stay monotone whatever order the chunks land in.  No workgroup
// ever waits on another one, so the kernels make progress whatever subset of
// their grid is resident -- e.g. while other streams' kernels share the CUs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gg {

// tile_start[t] = slot for every tile boundary t*T in [eo, eo + d).
__device__ __forceinline__ void mark_tiles(uint32_t* ts, uint32_t T, uint32_t eo, uint32_t d, uint32_t slot) {
    const uint64_t stop = uint64_t(eo) + d;
    for (uint64_t tt = (uint64_t(eo) + T - 1) / T; tt * T < stop; ++tt) ts[tt] = slot;
}

// Last entry of the tile whose edge prefix is <= e (binary search in LDS).
__device__ __forceinline__ uint32_t owner_of(const uint32_t* eoff, uint32_t c, uint32_t e) {
    uint32_t lo = 0, hi = c - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (eoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// True when the chunk reservation [base, base + total) stays inside a queue of
// n_cap entries and an edge range of m_cap edges.
__device__ __forceinline__ bool reservation_fits(unsigned long long base, unsigned long long total, uint64_t n_cap,
                                                 uint64_t m_cap) {
    const uint64_t cnt_end = (base >> 32) + (total >> 32);
    const uint64_t edge_end = (base & 0xffffffffull) + (total & 0xffffffffull);
    return cnt_end <= n_cap && edge_end <= m_cap;
}

}  // namespace gg
