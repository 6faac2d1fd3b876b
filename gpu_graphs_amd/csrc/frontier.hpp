// Device helpers shared by the single-source and multi-source BFS kernels:
// the decoupled look-back that turns per-chunk (count, edges) aggregates into
// global exclusive prefixes inside one launch, and the tile index that maps
// an edge offset to the frontier entry owning it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gg {

constexpr unsigned long long kFlagAgg = 1ull << 62;     // look-back: chunk aggregate published
constexpr unsigned long long kFlagPrefix = 2ull << 62;  // look-back: inclusive prefix published
constexpr unsigned long long kValueMask = (1ull << 62) - 1;

// Coherent read of a look-back status word.  Relaxed on purpose: the status
// words carry all the data exchanged inside a launch (the queue they index
// is read by the next launch), and an agent-scope acquire would invalidate
// the whole L2 of the XCD on every poll (buffer_inv sc1 on gfx950).
__device__ __forceinline__ unsigned long long rmw_read(unsigned long long* p) {
    return __hip_atomic_fetch_add(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back (Merrill & Garland 2016) for chunk c, called by one
// thread: publishes the chunk's aggregate, sums predecessors until one with
// an inclusive prefix, publishes its own inclusive prefix and returns the
// exclusive one.  Values are (count << 32 | edges), at most 62 bits.  The
// wait is bounded by `spin_limit` wall-clock ticks; on timeout `err_bit` is
// set in *err and the partial sum is returned.  Chunks are assigned to
// workgroups statically and in order, and every workgroup of the launch is
// resident, so a predecessor is always making progress.
__device__ inline unsigned long long lookback(unsigned long long* st, uint32_t c, unsigned long long agg,
                                              uint64_t spin_limit, unsigned long long* err,
                                              unsigned long long err_bit) {
    if (c == 0) {
        atomicExch(&st[0], kFlagPrefix | agg);
        return 0;
    }
    atomicExch(&st[c], kFlagAgg | agg);
    unsigned long long excl = 0;
    int64_t j = static_cast<int64_t>(c) - 1;
    const uint64_t t0 = wall_clock64();
    while (j >= 0) {
        const unsigned long long sj = rmw_read(&st[j]);
        const unsigned long long flag = sj & ~kValueMask;
        if (flag == kFlagPrefix) {
            excl += sj & kValueMask;
            break;
        }
        if (flag == kFlagAgg) {
            excl += sj & kValueMask;
            --j;
            continue;
        }
        if (static_cast<uint64_t>(wall_clock64()) - t0 > spin_limit) {
            atomicOr(err, err_bit);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    atomicExch(&st[c], kFlagPrefix | (excl + agg));
    return excl;
}

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

}  // namespace gg
