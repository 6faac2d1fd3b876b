// gpugraphs.hpp -- header-only C++ facade over gb.h with the reference's
// class and member names, so code written against YKMorsy/GPU_Graphs'
// `csr` / `gpuBFS` (1D Partition/include/csr.h:9-23,
// 1D Partition/include/gpuBFS.cuh:9-54) compiles against this engine:
//
//     gpugraphs::csr graph(file);               // csr::csr(const char*)
//     gpugraphs::gpuBFS bfs(graph, source);     // gpuBFS::gpuBFS(csr&, int)
//     bfs.host_distance[i]; bfs.exec_time;      // same members
//
// Differences: row_offset is int64_t, unreached vertices are -1 (not
// INT_MAX), and errors throw gpugraphs::error instead of exit().
#ifndef GPUGRAPHS_HPP
#define GPUGRAPHS_HPP

#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gb.h"

namespace gpugraphs {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int status) {
    if (status != GB_OK) throw error(status, gb_last_error());
}

class csr {
   public:
    explicit csr(const char* filename, uint32_t flags = 0) {
        check(gb_graph_load_edgelist(filename, flags, &g_));
        bind();
    }
    explicit csr(gb_graph* owned) : g_(owned) { bind(); }
    ~csr() {
        if (dg_) gb_device_free(dg_);
        if (g_) gb_graph_free(g_);
    }
    csr(const csr&) = delete;
    csr& operator=(const csr&) = delete;

    void print_info() const {
        std::cout << "\n------CSR INFO------" << std::endl;
        std::cout << "Column indices: " << num_edges << std::endl;
        for (int64_t i = 0; i < num_edges; i++) std::cout << col_idx[i] << " | ";
        std::cout << std::endl << "Row offset: " << num_nodes << std::endl;
        for (int64_t i = 0; i < num_nodes + 1; i++) std::cout << row_offset[i] << " | ";
        std::cout << std::endl;
    }

    // Device copy, uploaded on first use and kept resident.
    gb_dgraph* device(int dev = 0) {
        if (!dg_) check(gb_device_upload(g_, dev, &dg_));
        return dg_;
    }
    gb_graph* handle() const { return g_; }

    const int32_t* col_idx = nullptr;
    const int64_t* row_offset = nullptr;
    int64_t num_nodes = 0;
    int64_t num_edges = 0;

   private:
    void bind() {
        check(gb_graph_info(g_, &num_nodes, &num_edges));
        check(gb_graph_csr(g_, &row_offset, &col_idx));
    }
    gb_graph* g_ = nullptr;
    gb_dgraph* dg_ = nullptr;
};

class gpuBFS {
   public:
    gpuBFS(csr& graph, int source, int device = 0) : distance_(static_cast<size_t>(graph.num_nodes)) {
        gb_stats st{};
        check(gb_bfs(graph.device(device), source, distance_.data(), nullptr, &st));
        host_distance = distance_.data();
        exec_time = static_cast<float>(st.ms);
        copy_time = static_cast<float>(st.copy_ms);
        iteration = st.num_levels;
        total_edges_traversed = st.m_reached;
    }
    void print_distance(const csr& graph) const {
        std::cout << "\n------GPU DISTANCE VECTOR------" << std::endl;
        for (int64_t i = 0; i < graph.num_nodes; i++) std::cout << host_distance[i] << " | ";
        std::cout << std::endl;
    }

    int32_t* host_distance = nullptr;
    float exec_time = 0;
    float copy_time = 0;  // ms of the device-to-host copy of host_distance (get_device_distance)
    int64_t total_edges_traversed = 0;
    int iteration = 0;

   private:
    std::vector<int32_t> distance_;
};

}  // namespace gpugraphs

#endif
