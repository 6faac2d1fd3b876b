// graphBFS -- command-line surface of the reference (1D Partition/main.cu:
// 10-82): `graphBFS <input_file> <source_num>` loads the edge list, runs the
// BFS on the GPU and prints "<device> GPU BFS Time: X ms" (main.cu:35).
//
// Extra flags: --one-based, --symmetrize, --keep-order, --print (distance
// vector, as print_distance), --json (one JSON line with the statistics).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "gpugraphs/gpugraphs.hpp"

int main(int argc, char* argv[]) {
    const char* file = nullptr;
    const char* src_arg = nullptr;
    uint32_t flags = 0;
    bool print = false, json = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--one-based") flags |= GB_ONE_BASED;
        else if (a == "--symmetrize") flags |= GB_SYMMETRIZE;
        else if (a == "--keep-order") flags |= GB_KEEP_ORDER;
        else if (a == "--print") print = true;
        else if (a == "--json") json = true;
        else if (!file) file = argv[i];
        else if (!src_arg) src_arg = argv[i];
        else {
            file = nullptr;
            break;
        }
    }
    if (!file || !src_arg) {
        std::fprintf(stderr, "Usage: %s <input_file> <source_num> [--one-based] [--symmetrize] [--keep-order] "
                             "[--print] [--json]\n", argv[0]);
        return 1;
    }
    try {
        int source = std::stoi(src_arg);
        gpugraphs::csr graph(file, flags);
        gpugraphs::gpuBFS bfs(graph, source);
        if (print) bfs.print_distance(graph);
        if (json) {
            std::printf("{\"file\": \"%s\", \"source\": %d, \"n\": %lld, \"m\": %lld, \"levels\": %d, "
                        "\"edges_traversed\": %lld, \"ms\": %.4f, \"copy_ms\": %.4f}\n",
                        file, source, static_cast<long long>(graph.num_nodes), static_cast<long long>(graph.num_edges),
                        bfs.iteration, static_cast<long long>(bfs.total_edges_traversed), bfs.exec_time,
                        bfs.copy_time);
        } else {
            std::printf("0 GPU BFS Time: %g ms\n", bfs.exec_time);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "graphBFS: %s\n", e.what());
        return 2;
    }
    return 0;
}
