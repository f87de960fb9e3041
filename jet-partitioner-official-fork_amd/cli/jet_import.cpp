// Refinement-only replay with the reference's command line
// (app/import_coarse.cpp:149-204):
//   jet_import <config_file> [partition_output_file] [metrics_file]
// Reads coarse_graphs.out and coarse_part.out (written by jet_export, or by a
// partition call with dump_coarse set) from the working directory, refines the
// hierarchy on the GPU num_iter times and keeps the best cut by strict <.
#include <cstdio>
#include <vector>

#include "jet_capi.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "Insufficient number of args provided\n");
        std::fprintf(stderr, "Usage: %s <config_file> <optional partition_output_filename> <optional metrics_filename>\n",
                     argv[0]);
        return -1;
    }
    jet_config_t cfg;
    if (jet_load_config(argv[1], &cfg) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        return -1;
    }
    jet_hierarchy_t* h = nullptr;
    if (jet_load_hierarchy("coarse_graphs.out", JET_DUMP_32, &h) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        return -1;
    }
    const int32_t levels = jet_hierarchy_levels(h);
    std::printf("Importing %d coarsened graphs\n", levels);
    std::vector<int32_t> coarse((size_t)jet_hierarchy_vertices(h, levels - 1));
    if (jet_load_coarse_part("coarse_part.out", (int64_t)coarse.size(), coarse.data()) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        jet_free_hierarchy(h);
        return -1;
    }
    const int64_t n = jet_hierarchy_vertices(h, 0);
    FILE* metrics = argc >= 4 ? std::fopen(argv[3], "w") : nullptr;
    std::vector<int32_t> part((size_t)n), best((size_t)n);
    long long best_cut = -1;
    for (int i = 0; i < cfg.num_iter; i++) {
        int64_t cut = 0;
        jet_stats_t st;
        if (jet_uncoarsen_hierarchy(&cfg, h, coarse.data(), part.data(), &cut, &st) != JET_OK) {
            std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
            jet_free_hierarchy(h);
            return -1;
        }
        if (best_cut < 0 || cut < best_cut) {
            best_cut = cut;
            best.swap(part);
        }
        if (metrics)
            std::fprintf(metrics,
                         "{\"attempt\": %d, \"edge_cut\": %lld, \"total_s\": %.6f, \"refine_s\": %.6f, "
                         "\"levels\": %d, \"imbalance\": %.6f}\n",
                         i, (long long)cut, st.total_s, st.refine_s, st.num_levels, st.finest_imb_ratio);
    }
    if (metrics) std::fclose(metrics);
    std::printf("Imported coarse graphs, min edgecut found is %lld\n", best_cut);
    if (argc >= 3 && cfg.num_iter > 0 && jet_write_partition(argv[2], n, best.data()) != JET_OK)
        std::fprintf(stderr, "could not write %s: %s\n", argv[2], jet_last_error());
    jet_free_hierarchy(h);
    return 0;
}
