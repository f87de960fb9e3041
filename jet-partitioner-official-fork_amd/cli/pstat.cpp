// Partition quality report with the reference's command line
// (app/part_eval.cpp:46-115):  pstat <metis_graph_file> <part_file> <k>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "jet_capi.h"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "Insufficient number of args provided\n");
        std::fprintf(stderr, "Usage: %s <metis_graph_file> <part_file> <k>\n", argv[0]);
        return -1;
    }
    const int k = std::atoi(argv[3]);
    jet_host_graph_t g;
    if (jet_load_metis(argv[1], &g) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        return -1;
    }
    std::printf("vertices: %lld; edges: %lld\n", (long long)g.n, (long long)(g.m / 2));
    std::vector<int32_t> part(g.n);
    jet_quality_t q;
    if (jet_read_partition(argv[2], g.n, part.data()) != JET_OK ||
        jet_partition_quality(g.n, g.row_offsets, g.col, g.ew, nullptr, part.data(), k, JET_PTR_HOST, &q) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        jet_free_graph(&g);
        return -1;
    }
    std::printf("Max part cut: %lld\n", (long long)q.max_part_cut);
    std::printf("Cutsize: %lld\n", (long long)q.edge_cut);
    std::printf("Comm size: %lld\n", (long long)q.comm_size);
    std::printf("Largest: %.5g\n", q.max_imb);
    std::printf("Smallest: %.5g\n", q.min_imb);
    jet_free_graph(&g);
    return 0;
}
