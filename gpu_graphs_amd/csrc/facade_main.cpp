// facade_main -- a caller written the way a user of the reference writes
// one: the flow of 1D Partition/main.cu:10-35 (usage check, `csr graph(file)`,
// `gpuBFS gpuBFS(graph, source)`, the "<pe> GPU BFS Time: X ms" line), built
// against this engine with only the include change shown in the integration
// notes, plus the distance vector the reference prints with
// gpuBFS::print_distance (its main has that call commented out).
#include <iostream>
#include <string>

#include "gpugraphs/gpugraphs.hpp"
using gpugraphs::csr;
using gpugraphs::gpuBFS;
static int nvshmem_my_pe() { return 0; }  // one process, one GPU

int main(int argc, char* argv[]) {
    if (argc != 3) {
        std::cerr << "Usage: " << argv[0] << " <input_file> <source_num>" << std::endl;
        return 1;
    }
    const char* file = argv[1];
    int source = std::stoi(argv[2]);
    csr graph(file);
    gpuBFS search(graph, source);
    std::cout << nvshmem_my_pe() << " GPU BFS Time: " << search.exec_time << " ms" << std::endl;
    search.print_distance(graph);
    return 0;
}
