// ref_driver.cpp -- C entry points around the reference's own CPU classes.
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile together with the
// unmodified reference sources `1D Partition/src/csr.cpp` and
// `1D Partition/src/cpuBFS.cpp` (compiled where they lie) into
// oracle/_ref/libref.so.  Tests use it to check the C restatement in this
// folder against the reference itself; bench.py may time it as the CPU
// baseline.  Nothing in the product loads it.
//
// The reference classes are used through their public members only:
//   csr::csr(const char*), row_offset, col_idx, num_nodes, num_edges
//                                            (1D Partition/include/csr.h:12-20)
//   cpuBFS::cpuBFS(csr&, int), distance, exec_time, iteration
//                                            (1D Partition/include/cpuBFS.h:16-24)
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <unistd.h>

#include "cpuBFS.h"
#include "csr.h"

namespace {

// A csr object is built from a file; an empty graph file gives one whose
// arrays can be replaced by an in-memory CSR (the destructor frees them).
std::string empty_graph_file() {
    static std::string path;
    if (path.empty()) {
        char tmpl[] = "/tmp/gb_ref_empty_XXXXXX";
        int fd = mkstemp(tmpl);
        if (fd < 0) return "";
        const char body[] = "0\n0\n";
        if (write(fd, body, sizeof(body) - 1) != static_cast<ssize_t>(sizeof(body) - 1)) {
            close(fd);
            return "";
        }
        close(fd);
        path = tmpl;
    }
    return path;
}

}  // namespace

extern "C" {

// Loads `path` with csr::csr.  Call with row_offset/col_idx NULL to get the
// sizes, then again with buffers of n+1 and m ints.  Returns 0 or -2.
int ref_load_csr(const char* path, int* n, int* m, int* row_offset, int* col_idx) {
    FILE* probe = std::fopen(path, "r");
    if (!probe) return -2;
    std::fclose(probe);
    csr g(path);
    *n = g.num_nodes;
    *m = g.row_offset[g.num_nodes];
    if (row_offset && col_idx) {
        std::memcpy(row_offset, g.row_offset, sizeof(int) * (g.num_nodes + 1));
        std::memcpy(col_idx, g.col_idx, sizeof(int) * (*m));
    }
    return 0;
}

// Runs cpuBFS::cpuBFS on an in-memory CSR (int offsets, as the reference
// uses).  Levels are converted to int32 with -1 for the reference's INT_MAX.
// *ms is the reference's own clock() measurement (cpuBFS.cpp:6-7,60-61);
// *wall_ms is steady wall time around the constructor.
int ref_cpu_bfs(int n, const int* row_offset, const int* col_idx, int source, int32_t* levels, float* ms,
                double* wall_ms, int* iterations) {
    std::string empty = empty_graph_file();
    if (empty.empty()) return -2;
    csr g(empty.c_str());
    std::free(g.col_idx);
    std::free(g.row_offset);
    const int m = row_offset[n];
    g.num_nodes = n;
    g.num_edges = m;
    g.row_offset = static_cast<int*>(std::malloc(sizeof(int) * (n + 1)));
    g.col_idx = static_cast<int*>(std::malloc(sizeof(int) * (m > 0 ? m : 1)));
    if (!g.row_offset || !g.col_idx) return -3;
    std::memcpy(g.row_offset, row_offset, sizeof(int) * (n + 1));
    std::memcpy(g.col_idx, col_idx, sizeof(int) * m);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    cpuBFS bfs(g, source);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (wall_ms) *wall_ms = (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6;
    if (ms) *ms = bfs.exec_time;
    if (iterations) *iterations = bfs.iteration;
    for (int i = 0; i < n; ++i) {
        long long d = bfs.distance[i];
        levels[i] = d == INT_MAX ? -1 : static_cast<int32_t>(d);
    }
    return 0;
}

}  // extern "C"
