// Shared host utilities: status-carrying exceptions, the thread-local error
// string behind gb_last_error(), and a small std::thread parallel-for.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpugraphs/gb.h"

namespace gg {

// Error carrying a GB_E* status.  Thrown inside the library, caught at the
// C ABI boundary (capi.cpp) and turned into a return code + message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

void set_last_error(const std::string& msg);
const char* last_error();

// Worker count for host-side graph construction.  GB_THREADS overrides; the
// default is the hardware concurrency capped at 16.
inline int host_threads() {
    if (const char* e = std::getenv("GB_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) return v;
    }
    unsigned hc = std::thread::hardware_concurrency();
    if (hc == 0) hc = 1;
    return hc > 16 ? 16 : static_cast<int>(hc);
}

// Runs fn(begin, end, worker) over [0, n) split into contiguous slices.
template <class F>
void parallel_for(int64_t n, F&& fn, int workers = 0) {
    if (workers <= 0) workers = host_threads();
    if (n < 4096 || workers == 1) {
        fn(int64_t(0), n, 0);
        return;
    }
    if (workers > n) workers = static_cast<int>(n);
    std::vector<std::thread> pool;
    pool.reserve(workers);
    for (int w = 0; w < workers; ++w) {
        int64_t b = n * w / workers, e = n * (w + 1) / workers;
        pool.emplace_back([&fn, b, e, w] { fn(b, e, w); });
    }
    for (auto& t : pool) t.join();
}

// Counter-based 64-bit mixer (SplitMix64 finaliser, Steele et al. 2014).
// Generators draw from mix64(seed, index) so results do not depend on the
// number of worker threads.
inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i * 0x9E3779B97F4A7C15ull);
}

}  // namespace gg
