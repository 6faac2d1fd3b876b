/*
 * csr.c -- edge-list loader oracle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates csr::csr (1D Partition/src/csr.cpp:3-67): read `num_nodes`,
 * `num_edges` (csr.cpp:16-17), then `from to` pairs until the stream fails
 * (csr.cpp:26-29); row i lists its targets in file order with duplicates
 * (csr.cpp:32-54); row_offset[n] is the number of entries (csr.cpp:57) and
 * empty rows take the next row's offset (csr.cpp:60-65), i.e. offsets are
 * the exclusive prefix sum of out-degrees.  The 1-based variant subtracts
 * one from both ids (MIG/include/csr/csr.cpp:28).
 *
 * Departures, which the reference leaves undefined: ids outside [0, n)
 * after the shift are rejected (-5) instead of overflowing adjList
 * (csr.cpp:28), and the header edge count is not trusted for sizing.
 */
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

int oracle_load_edgelist(const char* path, int one_based, int64_t* n_out, int64_t* m_out, int64_t* offsets,
                         int32_t* cols) {
    FILE* f = fopen(path, "r");
    if (!f) return -2;
    long long n = 0, m_hint = 0, a = 0, b = 0;
    if (fscanf(f, "%lld %lld", &n, &m_hint) != 2 || n < 0) {
        fclose(f);
        return -2;
    }
    const long long shift = one_based ? 1 : 0;
    int64_t* deg = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (!deg) {
        fclose(f);
        return -3;
    }
    int64_t m = 0;
    /* Pass 1: count per-row entries (the reference's push_back). */
    while (fscanf(f, "%lld %lld", &a, &b) == 2) {
        a -= shift;
        b -= shift;
        if (a < 0 || a >= n || b < 0 || b >= n) {
            free(deg);
            fclose(f);
            return -5;
        }
        deg[a]++;
        m++;
    }
    *n_out = n;
    *m_out = m;
    if (!offsets || !cols) {
        free(deg);
        fclose(f);
        return 0;
    }
    offsets[0] = 0;
    for (int64_t v = 0; v < n; ++v) offsets[v + 1] = offsets[v] + deg[v];
    /* Pass 2: place targets in file order. */
    for (int64_t v = 0; v < n; ++v) deg[v] = offsets[v];
    rewind(f);
    if (fscanf(f, "%lld %lld", &n, &m_hint) != 2) {
        free(deg);
        fclose(f);
        return -2;
    }
    while (fscanf(f, "%lld %lld", &a, &b) == 2) {
        a -= shift;
        b -= shift;
        cols[deg[a]++] = (int32_t)b;
    }
    free(deg);
    fclose(f);
    return 0;
}
