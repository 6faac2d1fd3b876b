// Device-wide exclusive scan of int32 (wrap-around arithmetic).
//
// Replaces PrefixSum::sum_scan_blelloch (1D Partition/src/prefixSum.cu:
// 139-224): 1024-element Blelloch tiles in padded shared memory, block sums
// scanned recursively, and cudaMalloc/cudaFree of temporaries on every call
// (prefixSum.cu:167-168,185-186,194,201,204,223).
//
// Here: reduce-then-scan over 4096-element tiles (256 threads x 16 items).
//   1. tile_reduce:  per-tile sum -> partial[tile]
//   2. partial_scan: one workgroup scans partial[] in place (exclusive)
//   3. tile_scan:    thread-serial scan of 16 items, wave64 scan of thread
//                    totals, LDS-staged scan of the 4 wave totals, + partial
// The partial array is cached per device and only grows, so steady-state
// calls allocate nothing.  exclusive_scan_i32_noalloc never allocates: its
// partials live in a fixed device array (`scratch`), which bounds its length.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "device_graph.hpp"
#include "wave.hpp"

namespace gg {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
// LDS row padding: one extra word every 32 words breaks the 16-word stride
// of the blocked layout across LDS banks.
__device__ __forceinline__ uint32_t pad(uint32_t i) { return i + (i >> 5); }

__global__ __launch_bounds__(kBlock) void tile_reduce(const uint32_t* __restrict__ in, uint32_t* partial, int64_t n) {
    __shared__ uint32_t s_wave[kBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        int64_t i = base + j * kBlock + threadIdx.x;
        if (i < n) sum += in[i];
    }
    sum = wave_sum(sum);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += s_wave[w];
        partial[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void partial_scan(uint32_t* partial, int64_t count) {
    __shared__ uint32_t s_wave[1024 / 64];
    uint32_t carry = 0;
    for (int64_t b = 0; b < count; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const uint32_t x = i < count ? partial[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<1024>(x, s_wave, tot);
        if (i < count) partial[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kBlock) void tile_scan(const uint32_t* in, uint32_t* out, const uint32_t* partial,
                                                     int64_t n) {
    __shared__ uint32_t s_data[kTile + kTile / 32];
    __shared__ uint32_t s_wave[kBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    // Coalesced load into LDS.
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint32_t li = j * kBlock + threadIdx.x;
        const int64_t i = base + li;
        s_data[pad(li)] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    // Thread-serial exclusive scan of 16 consecutive items.
    uint32_t v[kItems];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        v[j] = s_data[pad(threadIdx.x * kItems + j)];
        const uint32_t t = v[j];
        v[j] = run;
        run += t;
    }
    uint32_t tot;
    const uint32_t before = block_exclusive_scan<kBlock>(run, s_wave, tot) + partial[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kItems; ++j) s_data[pad(threadIdx.x * kItems + j)] = v[j] + before;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint32_t li = j * kBlock + threadIdx.x;
        const int64_t i = base + li;
        if (i < n) out[i] = s_data[pad(li)];
    }
}

struct ScanCache {
    std::mutex mu;
    std::vector<uint32_t*> partial;
    std::vector<int64_t> cap;
};
ScanCache& cache() {
    static ScanCache c;
    return c;
}

// The three launches of one scan on the legacy default stream (ordered after
// the caller's copies), then a wait for them.
void run_scan(const int32_t* d_in, int32_t* d_out, int64_t n, int64_t tiles, uint32_t* partial) {
    const uint32_t* in = reinterpret_cast<const uint32_t*>(d_in);
    uint32_t* out = reinterpret_cast<uint32_t*>(d_out);
    hipStream_t s = nullptr;
    hipLaunchKernelGGL(tile_reduce, dim3(static_cast<uint32_t>(tiles)), dim3(kBlock), 0, s, in, partial, n);
    hipLaunchKernelGGL(partial_scan, dim3(1), dim3(1024), 0, s, partial, tiles);
    hipLaunchKernelGGL(tile_scan, dim3(static_cast<uint32_t>(tiles)), dim3(kBlock), 0, s, in, out, partial, n);
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(s));
}

}  // namespace

// Partials of the allocation-free scan: one per 4096-item tile, so inputs of
// up to 2^18 tiles = 2^30 items (1 MiB of device memory, held by every GPU
// the library runs on until the process ends).  Host code never passes
// this name to a kernel: it stands for a device address only on the device;
// the host uses the address hipGetSymbolAddress returns for each device.
constexpr int64_t kNoAllocMaxTiles = int64_t(1) << 18;
static __device__ __align__(256) uint32_t scratch[kNoAllocMaxTiles];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on `device` (current), looked up once per
// device with hipGetSymbolAddress; fixed-size table, no allocation.
uint32_t* scratch_on(int device) {
    static std::mutex mu;
    static uint32_t* ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) fail(GB_EINVAL, "device id outside [0, 64)");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(GB_EDEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        ptr[device] = static_cast<uint32_t*>(p);
    }
    return ptr[device];
}

}  // namespace

int64_t exclusive_scan_noalloc_max() { return kNoAllocMaxTiles * kTile; }

void exclusive_scan_i32_noalloc(int device, const int32_t* d_in, int32_t* d_out, int64_t n) {
    if (n < 0) fail(GB_EINVAL, "negative scan length");
    if (n > kNoAllocMaxTiles * kTile)
        fail(GB_EINVAL, "scan of " + std::to_string(n) + " items exceeds the allocation-free maximum of " +
                            std::to_string(kNoAllocMaxTiles * kTile) + " (use gb_exclusive_scan_i32)");
    if (n == 0) return;
    if (!d_in || !d_out) fail(GB_EINVAL, "null scan buffer");
    DeviceGuard guard(device);
    // Every partial the scan reads, tile_reduce writes first in the same
    // call: nothing from an earlier call needs clearing.
    run_scan(d_in, d_out, n, (n + kTile - 1) / kTile, scratch_on(device));
}

void exclusive_scan_i32(int device, const int32_t* d_in, int32_t* d_out, int64_t n) {
    if (n < 0) fail(GB_EINVAL, "negative scan length");
    if (n == 0) return;
    if (!d_in || !d_out) fail(GB_EINVAL, "null scan buffer");
    DeviceGuard guard(device);
    const int64_t tiles = (n + kTile - 1) / kTile;
    if (tiles > INT32_MAX) fail(GB_EINVAL, "scan too long");
    uint32_t* partial = nullptr;
    {
        ScanCache& c = cache();
        std::lock_guard<std::mutex> lk(c.mu);
        if (static_cast<int>(c.partial.size()) <= device) {
            c.partial.resize(device + 1, nullptr);
            c.cap.resize(device + 1, 0);
        }
        if (c.cap[device] < tiles) {
            if (c.partial[device]) GG_HIP(hipFree(c.partial[device]));
            c.partial[device] = nullptr;
            c.cap[device] = 0;
            GG_HIP(hipMalloc(&c.partial[device], sizeof(uint32_t) * tiles));
            c.cap[device] = tiles;
        }
        partial = c.partial[device];
    }
    run_scan(d_in, d_out, n, tiles, partial);
}

}  // namespace gg
