/*
 * bfs.c -- CPU BFS oracle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates cpuBFS::cpuBFS (1D Partition/src/cpuBFS.cpp:5-62): a serial,
 * level-synchronous FIFO BFS.  The source gets level 0 (cpuBFS.cpp:29); a
 * neighbour whose level is still unset when it is first scanned gets
 * `iteration + 1` and is appended to the queue (cpuBFS.cpp:49-53); the
 * iteration counter advances once per level (cpuBFS.cpp:57).  The reference
 * keeps a std::queue and a `level_size` loop; a flat array with a head index
 * and a per-level end index visits vertices in the same order.
 */
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static int bad_args(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source) {
    if (n < 1 || !offsets || (offsets[n] > 0 && !cols)) return -1;
    if (source < 0 || source >= n) return -5;
    return 0;
}

int oracle_bfs(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source, int32_t* levels,
               int32_t* parents, int64_t* n_reached, int64_t* m_reached, int32_t* num_levels) {
    int rc = bad_args(n, offsets, cols, source);
    if (rc) return rc;
    int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!queue) return -3;
    for (int64_t i = 0; i < n; ++i) levels[i] = -1;
    if (parents)
        for (int64_t i = 0; i < n; ++i) parents[i] = -1;
    int64_t head = 0, tail = 0, mr = 0;
    levels[source] = 0;
    if (parents) parents[source] = source;
    queue[tail++] = source;
    int32_t iteration = 0;
    while (head < tail) {
        const int64_t level_end = tail;
        for (; head < level_end; ++head) {
            const int32_t v = queue[head];
            mr += offsets[v + 1] - offsets[v];
            for (int64_t j = offsets[v]; j < offsets[v + 1]; ++j) {
                const int32_t nb = cols[j];
                if (levels[nb] == -1) {
                    levels[nb] = iteration + 1;
                    if (parents) parents[nb] = v;
                    queue[tail++] = nb;
                }
            }
        }
        ++iteration;
    }
    free(queue);
    if (n_reached) *n_reached = tail;
    if (m_reached) *m_reached = mr;
    if (num_levels) *num_levels = iteration;
    return 0;
}

int oracle_bfs_timed(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source, int32_t* levels_out,
                     double* ms, int64_t* n_reached, int64_t* m_reached) {
    int rc = bad_args(n, offsets, cols, source);
    if (rc) return rc;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int32_t* levels = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!levels) return -3;
    rc = oracle_bfs(n, offsets, cols, source, levels, NULL, n_reached, m_reached, NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (ms) *ms = (double)(t1.tv_sec - t0.tv_sec) * 1e3 + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-6;
    if (levels_out && rc == 0) memcpy(levels_out, levels, sizeof(int32_t) * (size_t)n);
    free(levels);
    return rc;
}

/* cpuBFS::cpuBFSAlt (cpuBFS.cpp:64-106): every neighbour of a dequeued,
 * first-time-seen vertex is enqueued unconditionally; a vertex takes the
 * current iteration as its level when it is dequeued for the first time.
 * At most 1 + sum of degrees of reached vertices entries are ever queued. */
int oracle_bfs_dequeue(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source, int32_t* levels) {
    int rc = bad_args(n, offsets, cols, source);
    if (rc) return rc;
    const int64_t cap = offsets[n] + 1;
    int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    if (!queue) return -3;
    for (int64_t i = 0; i < n; ++i) levels[i] = -1;
    int64_t head = 0, tail = 0;
    queue[tail++] = source;
    int32_t iteration = 0;
    while (head < tail) {
        const int64_t level_end = tail;
        for (; head < level_end; ++head) {
            const int32_t v = queue[head];
            if (levels[v] != -1) continue;
            levels[v] = iteration;
            for (int64_t j = offsets[v]; j < offsets[v + 1]; ++j) queue[tail++] = cols[j];
        }
        ++iteration;
    }
    free(queue);
    return 0;
}

int oracle_bfs_multi(int64_t n, const int64_t* offsets, const int32_t* cols, const int32_t* sources, int k,
                     int32_t* levels) {
    for (int i = 0; i < k; ++i) {
        int rc = oracle_bfs(n, offsets, cols, sources[i], levels + (int64_t)i * n, NULL, NULL, NULL, NULL);
        if (rc) return rc;
    }
    return 0;
}
