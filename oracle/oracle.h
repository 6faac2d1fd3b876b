/*
 * oracle.h -- CPU restatement of the reference algorithms on the BFS path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpu_graphs_amd/,
 * include/) links, loads or calls this code.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it, and only as
 * the checker / the CPU timing baseline, never as a result that is shipped.
 *
 * Conventions: 0-based int32 vertex ids, int64 CSR offsets, levels int32 with
 * -1 for "not reached" (the reference's INT_MAX, 1D Partition/include/
 * cpuBFS.h:11).  Functions return 0 on success, a negative code on bad input.
 */
#ifndef GB_ORACLE_H
#define GB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* csr.c -- the reference loader csr::csr (1D Partition/src/csr.cpp:3-67):
 * file order and duplicates kept, optional 1-based ids (MIG/include/csr/
 * csr.cpp:28).  Two calls: first with offsets/cols NULL to get n and m. */
int oracle_load_edgelist(const char* path, int one_based, int64_t* n, int64_t* m,
                         int64_t* offsets, int32_t* cols);

/* bfs.c -- cpuBFS::cpuBFS (1D Partition/src/cpuBFS.cpp:5-62). */
int oracle_bfs(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
               int32_t* levels, int32_t* parents, int64_t* n_reached, int64_t* m_reached,
               int32_t* num_levels);
/* Same traversal, timed like the reference: the clock covers allocating and
 * initialising the level array as well as the traversal (cpuBFS.cpp:6-24). */
int oracle_bfs_timed(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                     int32_t* levels_out, double* ms, int64_t* n_reached, int64_t* m_reached);
/* cpuBFS::cpuBFSAlt (cpuBFS.cpp:64-106): visit-on-dequeue, the model of the
 * reference GPU kernel linear_bfs (gpuBFS_kernels.cu:35). */
int oracle_bfs_dequeue(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                       int32_t* levels);
/* k sources, levels row-major [k][n]. */
int oracle_bfs_multi(int64_t n, const int64_t* offsets, const int32_t* cols, const int32_t* sources,
                     int k, int32_t* levels);

/* validate.c -- level invariants and Graph500 BFS-tree rules. Returns 0 if
 * valid, else a positive code; *where = offending vertex. */
int oracle_check_levels(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                        const int32_t* levels, int64_t* where);
int oracle_check_parents(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                         const int32_t* levels, const int32_t* parents, int64_t* where);

/* scan.c -- PrefixSum::sum_scan_blelloch (1D Partition/src/prefixSum.cu:
 * 139-224) semantics: exclusive scan, 32-bit wrap-around. */
void oracle_exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n);
/* The Blelloch up-sweep / down-sweep on a power-of-two array, in place
 * (prefixSum.cu:91-136, misc/test_prefix.cu:17-74). */
int oracle_blelloch_scan_u32(uint32_t* data, int64_t n_pow2);

/* serialized.c -- the scan-compaction model serialized::serialized
 * (1D Partition/src/serialized.cpp:3-162).  Writes levels and, per level,
 * the frontier size; frontier_sizes needs room for n+1 entries. */
int oracle_serialized_bfs(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t source,
                          int32_t* levels, int64_t* frontier_sizes, int32_t* num_levels);

#ifdef __cplusplus
}
#endif
#endif
