/*
 * scan.c -- exclusive scan oracle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * PrefixSum::sum_scan_blelloch (1D Partition/src/prefixSum.cu:139-224)
 * produces the exclusive prefix sum d_out[i] = d_in[0] + ... + d_in[i-1] of
 * `int` values; overflow wraps in two's complement on the GPU, so the oracle
 * uses uint32 arithmetic.
 *
 * oracle_blelloch_scan_u32 restates the work-efficient tree of one scan tile
 * (prefixSum.cu:91-136; misc/test_prefix.cu:17-50, Blelloch 1990): an
 * up-sweep that leaves partial sums at the right child of every subtree, the
 * root cleared to 0, then a down-sweep that swaps and adds down the tree.
 */
#include "oracle.h"

void oracle_exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n) {
    uint32_t run = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
}

int oracle_blelloch_scan_u32(uint32_t* a, int64_t n) {
    if (n < 1 || (n & (n - 1))) return -1;
    int64_t offset = 1;
    for (int64_t d = n >> 1; d > 0; d >>= 1) { /* up-sweep */
        for (int64_t t = 0; t < d; ++t) {
            const int64_t ai = offset * (2 * t + 1) - 1, bi = offset * (2 * t + 2) - 1;
            a[bi] += a[ai];
        }
        offset <<= 1;
    }
    a[n - 1] = 0;
    for (int64_t d = 1; d < n; d <<= 1) { /* down-sweep */
        offset >>= 1;
        for (int64_t t = 0; t < d; ++t) {
            const int64_t ai = offset * (2 * t + 1) - 1, bi = offset * (2 * t + 2) - 1;
            const uint32_t tmp = a[ai];
            a[ai] = a[bi];
            a[bi] += tmp;
        }
    }
    return 0;
}
