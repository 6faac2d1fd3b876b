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

// Degree buckets: bucket 0 holds the vertices of degree 0, bucket b >= 1
// those with degree in [2^(b-1), 2^b).
constexpr int kDegreeBuckets = 34;

struct HostGraph {
    int64_t n = 0;
    std::vector<int64_t> offsets;  // n + 1
    std::vector<int32_t> cols;     // offsets[n]
    uint32_t flags = 0;            // GB_* flags it was built with
    bool symmetric = false;        // built with GB_SYMMETRIZE or by a generator
    // Degree-sorted vertex buckets (filled by compute_degree_buckets):
    // bucket b is bucket_vertices[bucket_offsets[b] .. bucket_offsets[b+1]),
    // ids ascending within a bucket.
    std::vector<int64_t> bucket_offsets;
    std::vector<int32_t> bucket_vertices;

    int64_t m() const { return offsets.empty() ? 0 : offsets.back(); }
    int64_t degree(int64_t v) const { return offsets[v + 1] - offsets[v]; }
};

// Bucket index of a degree (see kDegreeBuckets).
inline int degree_bucket(int64_t d) {
    if (d <= 0) return 0;
    int b = 1;
    while (b < kDegreeBuckets - 1 && (int64_t(1) << b) <= d) ++b;
    return b;
}

void compute_degree_buckets(HostGraph& g);

// True when every edge (u, v) has its reverse (v, u): built symmetric, or
// checked on sorted rows.  Graphs kept in file order (GB_KEEP_ORDER) that were
// not built symmetric report false.
bool is_symmetric(const HostGraph& g);

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
