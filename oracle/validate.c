/*
 * validate.c -- BFS result checks.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference validates by counting mismatches against cpuBFS (the loop is
 * commented out at 1D Partition/main.cu:45-75).  These checks pin a level
 * array without another BFS: a labelling is the BFS distance from `source`
 * exactly when
 *   (1) level[source] = 0,
 *   (2) every edge u->v with u reached has v reached and
 *       level[v] <= level[u] + 1,
 *   (3) every reached v != source has an in-neighbour u with
 *       level[u] = level[v] - 1.
 * (2) bounds levels from above by the distance, (3) from below.  Parents
 * follow the Graph500 BFS-tree rules: parent[source] = source,
 * parent[v] = -1 iff level[v] = -1, (parent[v], v) is an edge and
 * level[parent[v]] = level[v] - 1.
 */
#include <stdlib.h>

#include "oracle.h"

static int has_edge(const int64_t* offsets, const int32_t* cols, int32_t u, int32_t v) {
    for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j)
        if (cols[j] == v) return 1;
    return 0;
}

int oracle_check_levels(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                        const int32_t* levels, int64_t* where) {
    if (source < 0 || source >= n) return -5;
    *where = source;
    if (levels[source] != 0) return 1;
    unsigned char* supported = (unsigned char*)calloc((size_t)n, 1);
    if (!supported) return -3;
    supported[source] = 1;
    for (int64_t u = 0; u < n; ++u) {
        const int32_t lu = levels[u];
        if (lu < -1) {
            *where = u;
            free(supported);
            return 2;
        }
        if (lu < 0) continue;
        for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j) {
            const int32_t v = cols[j];
            if (levels[v] < 0 || levels[v] > lu + 1) {
                *where = v;
                free(supported);
                return 3;
            }
            if (levels[v] == lu + 1) supported[v] = 1;
        }
    }
    for (int64_t v = 0; v < n; ++v)
        if (levels[v] >= 0 && !supported[v]) {
            *where = v;
            free(supported);
            return 4;
        }
    free(supported);
    return 0;
}

int oracle_check_parents(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                         const int32_t* levels, const int32_t* parents, int64_t* where) {
    if (source < 0 || source >= n) return -5;
    *where = source;
    if (parents[source] != source || levels[source] != 0) return 1;
    for (int64_t v = 0; v < n; ++v) {
        if (v == source) continue;
        *where = v;
        const int32_t p = parents[v];
        if ((p == -1) != (levels[v] == -1)) return 2;
        if (p == -1) continue;
        if (p < 0 || p >= n) return 3;
        if (levels[p] != levels[v] - 1) return 4;
        if (!has_edge(offsets, cols, p, (int32_t)v)) return 5;
    }
    return 0;
}
