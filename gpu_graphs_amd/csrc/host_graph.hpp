// Host CSR graph and its builders.
//
// Replaces csr::csr (1D Partition/src/csr.cpp:3-67, type csr.h:9-23).  The
// reference keeps `int row_offset[n+1]`, `int col_idx[m]` in file order with
// duplicates; this CSR uses int64 offsets and int32 columns, rows sorted and
// deduplicated by default (GB_KEEP_ORDER reproduces the reference layout).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gg {

struct HostGraph {
    int64_t n = 0;
    std::vector<int64_t> offsets;  // n + 1
    std::vector<int32_t> cols;     // offsets[n]
    uint32_t flags = 0;            // GB_* flags it was built with
    bool symmetric = false;        // built with GB_SYMMETRIZE or by a generator

    int64_t m() const { return offsets.empty() ? 0 : offsets.back(); }
    int64_t degree(int64_t v) const { return offsets[v + 1] - offsets[v]; }
};

// Edge list -> CSR.  src/dst hold m directed edges; ids must lie in [0,n).
// flags: GB_SYMMETRIZE, GB_KEEP_ORDER, GB_DROP_SELF_LOOPS.
HostGraph build_csr(int64_t n, const int32_t* src, const int32_t* dst, int64_t m, uint32_t flags);

// Sort every row and drop duplicate columns (and self-loops if asked), in place.
void sort_dedup_rows(HostGraph& g, bool drop_self_loops);

// Reference text format (csr.cpp:16-29): `n`, `m`, then `from to` pairs.
HostGraph load_edgelist(const std::string& path, uint32_t flags);
void write_edgelist(const HostGraph& g, const std::string& path);

void save_binary(const HostGraph& g, const std::string& path);
HostGraph load_binary(const std::string& path);

// Generators (generators.cpp).
HostGraph generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                        bool permute_ids);
HostGraph generate_er(int64_t n_candidates, int64_t n_isolated, int64_t m_undirected, uint64_t seed);
HostGraph generate_grid(int64_t width, int64_t height, double extra_fraction, uint64_t seed);

}  // namespace gg
