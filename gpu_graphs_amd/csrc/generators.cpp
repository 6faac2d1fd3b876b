// Synthetic graph generators for the measured workloads.  The reference ships
// only small hand-made edge lists (graphs/*.txt); these produce the R-MAT,
// Erdos-Renyi and road-like grid inputs at full size, deterministically for a
// seed and independently of the host thread count (counter-based RNG).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <unordered_set>

#include "common.hpp"
#include "host_graph.hpp"

namespace gg {

// Recursive-matrix generator (Chakrabarti, Zhan, Faloutsos 2004): each of the
// edge_factor * 2^scale edges picks one of four quadrants per bit level with
// probabilities a, b, c, d = 1-a-b-c.  No noise is added.  Vertex ids are
// optionally shuffled by a seeded permutation (Graph500 practice), then the
// edge list is symmetrised with self-loops and duplicates removed.
HostGraph generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                        bool permute_ids) {
    if (scale < 1 || scale > 30) fail(GB_EINVAL, "R-MAT scale must lie in [1, 30]");
    if (edge_factor < 1) fail(GB_EINVAL, "R-MAT edge factor must be >= 1");
    double d = 1.0 - a - b - c;
    if (a < 0 || b < 0 || c < 0 || d < -1e-12) fail(GB_EINVAL, "R-MAT probabilities must be >= 0 and sum to <= 1");
    const int64_t n = int64_t(1) << scale;
    const int64_t m = int64_t(edge_factor) << scale;
    const uint64_t ta = static_cast<uint64_t>(a * 4294967296.0);
    const uint64_t tab = static_cast<uint64_t>((a + b) * 4294967296.0);
    const uint64_t tabc = static_cast<uint64_t>((a + b + c) * 4294967296.0);

    std::vector<int32_t> src(m), dst(m);
    parallel_for(m, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t x = rng_at(seed, 0, static_cast<uint64_t>(i));
            uint32_t u = 0, v = 0;
            for (int l = 0; l < scale; ++l) {
                if (l & 1) x = mix64(x);
                uint64_t r = (l & 1) ? (x >> 32) : (x & 0xffffffffu);
                uint32_t ub = r >= tab;                  // quadrants c, d: lower half
                uint32_t vb = (r >= ta && r < tab) || r >= tabc;  // quadrants b, d: right half
                u = (u << 1) | ub;
                v = (v << 1) | vb;
            }
            src[i] = static_cast<int32_t>(u);
            dst[i] = static_cast<int32_t>(v);
        }
    });
    if (permute_ids) {
        std::vector<int32_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        for (int64_t i = n - 1; i > 0; --i) {
            int64_t j = static_cast<int64_t>(rng_at(seed, 1, static_cast<uint64_t>(i)) % static_cast<uint64_t>(i + 1));
            std::swap(perm[i], perm[j]);
        }
        parallel_for(m, [&](int64_t lo, int64_t hi, int) {
            for (int64_t i = lo; i < hi; ++i) {
                src[i] = perm[src[i]];
                dst[i] = perm[dst[i]];
            }
        });
    }
    return build_csr(n, src.data(), dst.data(), m, GB_SYMMETRIZE | GB_DROP_SELF_LOOPS);
}

// Erdos-Renyi G(n, m) over the first n_candidates ids: m distinct unordered
// pairs {u, v}, u != v, drawn uniformly with rejection of repeats.  The last
// n_isolated ids get no edges.
HostGraph generate_er(int64_t n_candidates, int64_t n_isolated, int64_t m_undirected, uint64_t seed) {
    if (n_candidates < 2 || n_isolated < 0 || m_undirected < 0) fail(GB_EINVAL, "bad Erdos-Renyi size");
    const int64_t n = n_candidates + n_isolated;
    if (n > INT32_MAX) fail(GB_EINVAL, "Erdos-Renyi graph too large");
    if (m_undirected > n_candidates * (n_candidates - 1) / 2) fail(GB_EINVAL, "more edges than vertex pairs");
    std::unordered_set<uint64_t> seen;
    seen.reserve(static_cast<size_t>(m_undirected) * 2);
    std::vector<int32_t> src, dst;
    src.reserve(m_undirected);
    dst.reserve(m_undirected);
    uint64_t draw = 0;
    while (static_cast<int64_t>(src.size()) < m_undirected) {
        uint64_t r = rng_at(seed, 4, draw++);
        int64_t u = static_cast<int64_t>((r & 0xffffffffu) % static_cast<uint64_t>(n_candidates));
        int64_t v = static_cast<int64_t>((r >> 32) % static_cast<uint64_t>(n_candidates));
        if (u == v) continue;
        uint64_t key = static_cast<uint64_t>(std::min(u, v)) * static_cast<uint64_t>(n) + static_cast<uint64_t>(std::max(u, v));
        if (!seen.insert(key).second) continue;
        src.push_back(static_cast<int32_t>(u));
        dst.push_back(static_cast<int32_t>(v));
    }
    return build_csr(n, src.data(), dst.data(), m_undirected, GB_SYMMETRIZE | GB_DROP_SELF_LOOPS);
}

// Road-like grid: the minimum spanning tree of a width x height 4-neighbour
// grid under seeded random edge weights (unique: weight ties are broken by
// edge id), plus round(extra_fraction * non-tree edges) of the remaining grid
// edges chosen uniformly by seed.  The MST is found with Boruvka's algorithm,
// whose rounds are parallel passes over the edge array; with unique weights
// it equals the random-weight Kruskal tree.
HostGraph generate_grid(int64_t width, int64_t height, double extra_fraction, uint64_t seed) {
    if (width < 1 || height < 1) fail(GB_EINVAL, "grid sides must be >= 1");
    if (extra_fraction < 0 || extra_fraction > 1) fail(GB_EINVAL, "extra_fraction must lie in [0, 1]");
    const int64_t n = width * height;
    if (n > INT32_MAX) fail(GB_EINVAL, "grid too large");
    const int64_t eh = height * (width - 1);       // horizontal edges
    const int64_t E = eh + (height - 1) * width;   // + vertical edges
    if (E >= (int64_t(1) << 32)) fail(GB_EINVAL, "grid has too many edges");
    auto endpoints = [&](int64_t e, int32_t& u, int32_t& v) {
        if (e < eh) {
            int64_t y = e / (width - 1), x = e % (width - 1);
            u = static_cast<int32_t>(y * width + x);
            v = u + 1;
        } else {
            int64_t k = e - eh;  // vertical: (x, y) -- (x, y + 1)
            u = static_cast<int32_t>(k);
            v = static_cast<int32_t>(k + width);
        }
    };
    // key = random 32-bit weight in the high half, edge id in the low half.
    auto key_of = [&](int64_t e) -> uint64_t {
        return (rng_at(seed, 2, static_cast<uint64_t>(e)) & 0xffffffff00000000ull) | static_cast<uint64_t>(e);
    };
    const uint64_t kNone = ~0ull;

    std::vector<int32_t> comp(n);
    std::iota(comp.begin(), comp.end(), 0);
    std::vector<uint8_t> tree(E, 0);
    std::vector<uint64_t> best(n, kNone);
    std::vector<int32_t> parent(n);
    int64_t tree_edges = 0;
    for (;;) {
        // 1. Lightest edge leaving each component (atomic min on best[root]).
        std::atomic<bool> any{false};
        parallel_for(E, [&](int64_t lo, int64_t hi, int) {
            bool found = false;
            for (int64_t e = lo; e < hi; ++e) {
                int32_t u, v;
                endpoints(e, u, v);
                int32_t cu = comp[u], cv = comp[v];
                if (cu == cv) continue;
                found = true;
                uint64_t k = key_of(e);
                for (int32_t cc : {cu, cv}) {
                    uint64_t cur = __atomic_load_n(&best[cc], __ATOMIC_RELAXED);
                    while (k < cur && !__atomic_compare_exchange_n(&best[cc], &cur, k, true, __ATOMIC_RELAXED,
                                                                   __ATOMIC_RELAXED)) {
                    }
                }
            }
            if (found) any.store(true, std::memory_order_relaxed);
        });
        if (!any.load()) break;
        // 2. Hook every root along its lightest edge; mutual picks keep the
        //    smaller id as root.  Mark the chosen edges as tree edges.
        std::atomic<int64_t> added{0};
        parallel_for(n, [&](int64_t lo, int64_t hi, int) {
            int64_t local = 0;
            for (int64_t r = lo; r < hi; ++r) {
                parent[r] = static_cast<int32_t>(r);
                if (comp[r] != r || best[r] == kNone) continue;
                int64_t e = static_cast<int64_t>(best[r] & 0xffffffffu);
                int32_t u, v;
                endpoints(e, u, v);
                int32_t other = comp[u] == r ? comp[v] : comp[u];
                bool mutual = best[other] == best[r];
                if (mutual && r < other) continue;  // r stays root; the pair's edge is counted once by `other`
                parent[r] = other;
                tree[e] = 1;
                ++local;
            }
            added += local;
        });
        tree_edges += added.load();
        // 3. Pointer jumping over roots, then relabel every vertex.
        for (bool changed = true; changed;) {
            std::atomic<bool> ch{false};
            parallel_for(n, [&](int64_t lo, int64_t hi, int) {
                bool c = false;
                for (int64_t r = lo; r < hi; ++r) {
                    if (comp[r] != r) continue;
                    int32_t p = __atomic_load_n(&parent[r], __ATOMIC_RELAXED);
                    int32_t pp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
                    if (p != pp) {
                        __atomic_store_n(&parent[r], pp, __ATOMIC_RELAXED);
                        c = true;
                    }
                }
                if (c) ch.store(true, std::memory_order_relaxed);
            });
            changed = ch.load();
        }
        parallel_for(n, [&](int64_t lo, int64_t hi, int) {
            for (int64_t v = lo; v < hi; ++v) comp[v] = parent[comp[v]];
        });
        parallel_for(n, [&](int64_t lo, int64_t hi, int) {
            for (int64_t v = lo; v < hi; ++v) best[v] = kNone;
        });
    }
    if (tree_edges != n - 1) fail(GB_EINVAL, "grid spanning tree construction failed");
    best.clear();
    best.shrink_to_fit();
    parent.clear();
    parent.shrink_to_fit();
    comp.clear();
    comp.shrink_to_fit();

    // Extra edges: the k non-tree edges with the smallest seeded random keys.
    std::vector<uint64_t> cand;
    cand.reserve(E - tree_edges);
    for (int64_t e = 0; e < E; ++e)
        if (!tree[e]) cand.push_back((rng_at(seed, 3, static_cast<uint64_t>(e)) & 0xffffffff00000000ull) | static_cast<uint64_t>(e));
    int64_t k = static_cast<int64_t>(std::llround(extra_fraction * static_cast<double>(cand.size())));
    if (k > 0 && k < static_cast<int64_t>(cand.size())) std::nth_element(cand.begin(), cand.begin() + k, cand.end());
    for (int64_t i = 0; i < k; ++i) tree[cand[i] & 0xffffffffu] = 2;
    cand.clear();
    cand.shrink_to_fit();

    std::vector<int32_t> src, dst;
    src.reserve(tree_edges + k);
    dst.reserve(tree_edges + k);
    for (int64_t e = 0; e < E; ++e) {
        if (!tree[e]) continue;
        int32_t u, v;
        endpoints(e, u, v);
        src.push_back(u);
        dst.push_back(v);
    }
    return build_csr(n, src.data(), dst.data(), static_cast<int64_t>(src.size()), GB_SYMMETRIZE);
}

}  // namespace gg
