// Wave64 / workgroup primitives for gfx950 (CDNA4): lane masks from
// ballot + mbcnt, wave scans, and an LDS-staged block scan.
//
// These replace the reference's 32-lane shuffle scan `block_prefix_sum`
// (1D Partition/src/gpuBFS_kernels.cu:273-332) and its Blelloch LDS variant
// (1D Partition/src/syclBFS.cpp:14-63).  Written from the published methods
// (Blelloch 1990; Harris, Sengupta, Owens, GPU Gems 3 ch. 39) for 64-wide
// wavefronts: a Hillis-Steele scan inside the wave, the per-wave totals
// staged in LDS, one pass over the (at most 16) wave totals.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gg {

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of `mask` below the calling lane (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Orders this wave's LDS accesses: stores of every lane before the barrier
// are visible to loads of every lane after it (lanes of a wave run in
// lockstep; the fence stops the compiler moving LDS accesses across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= static_cast<uint32_t>(o)) x += y;
    }
    return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Exclusive scan over the workgroup.  `s_wave` must hold BLOCK/64 entries.
// Returns this thread's exclusive prefix and sets `total` (same in every
// thread).  Contains two __syncthreads(); every thread must call it.
template <int BLOCK, class T>
__device__ __forceinline__ T block_exclusive_scan(T x, T* s_wave, T& total) {
    constexpr int W = BLOCK / 64;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    T incl = wave_inclusive_scan(x);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    T before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        T s = s_wave[w];
        before += (static_cast<uint32_t>(w) < wave) ? s : T(0);
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + incl - x;
}

}  // namespace gg
