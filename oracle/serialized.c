/*
 * serialized.c -- scan-compaction BFS model.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates serialized::serialized (1D Partition/src/serialized.cpp:3-162),
 * the reference's CPU statement of a four-phase GPU level step:
 *   nextLayer    (serialized.cpp:72-93)   label every unseen neighbour
 *                iteration+1 and remember the CSR entry that saw it first;
 *   countDegrees (serialized.cpp:95-128)  per frontier slot, count the
 *                entries that own their child (and are not self-loops);
 *   blockPrefixSum (serialized.cpp:130-137) exclusive scan of those counts;
 *   gather       (serialized.cpp:139-162) write owned children to the next
 *                queue at the scanned offsets, in edge order.
 * The next frontier is therefore deterministic: ordered by (frontier slot,
 * CSR entry) of the discovering edge.  Used to pin per-level frontier sizes.
 */
#include <stdlib.h>

#include "oracle.h"

int oracle_serialized_bfs(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source, int32_t* levels,
                          int64_t* frontier_sizes, int32_t* num_levels) {
    if (n < 1 || source < 0 || source >= n) return -5;
    int64_t* parent_edge = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int32_t* in_q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* out_q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int64_t* scan = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    if (!parent_edge || !in_q || !out_q || !scan) {
        free(parent_edge);
        free(in_q);
        free(out_q);
        free(scan);
        return -3;
    }
    for (int64_t i = 0; i < n; ++i) {
        levels[i] = -1;
        parent_edge[i] = -1;
    }
    levels[source] = 0;
    in_q[0] = source;
    int64_t qsize = 1;
    int32_t iteration = 0;
    while (qsize) {
        frontier_sizes[iteration] = qsize;
        for (int64_t t = 0; t < qsize; ++t) { /* nextLayer */
            const int32_t u = in_q[t];
            for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j) {
                const int32_t v = cols[j];
                if (levels[v] == -1) {
                    parent_edge[v] = j;
                    levels[v] = iteration + 1;
                }
            }
        }
        scan[0] = 0; /* countDegrees + blockPrefixSum */
        for (int64_t t = 0; t < qsize; ++t) {
            const int32_t u = in_q[t];
            int64_t owned = 0;
            for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j)
                if (parent_edge[cols[j]] == j && cols[j] != u) ++owned;
            scan[t + 1] = scan[t] + owned;
        }
        for (int64_t t = 0; t < qsize; ++t) { /* gather */
            const int32_t u = in_q[t];
            int64_t pos = scan[t];
            for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j)
                if (parent_edge[cols[j]] == j && cols[j] != u) out_q[pos++] = cols[j];
        }
        qsize = scan[qsize];
        int32_t* tmp = in_q;
        in_q = out_q;
        out_q = tmp;
        ++iteration;
    }
    frontier_sizes[iteration] = 0;
    *num_levels = iteration;
    free(parent_edge);
    free(in_q);
    free(out_q);
    free(scan);
    return 0;
}
