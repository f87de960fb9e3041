// Partitioning driver with the reference's command line (app/driver.cpp:66-149):
//   jet <metis_graph_file> <config_file> [partition_output_file] [metrics_file]
// Builds: jet (refine_tolerance 0.999), jet4 (0.9999), jet2 (0.99), jet_host and
// jet_serial (the reference's host-memory variants; here every variant runs on
// the GPU), and jet_export (dump_coarse set: writes coarse_graphs.out and
// coarse_part.out for jet_import, with the reference driver's int32 graph
// type). Attempt i uses seed JET_SEED + i (default 0); the best cut is kept by
// strict <. The metrics file, when given, receives one JSON line per attempt.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "jet_capi.h"

#ifndef JET_REFINE_TOL
#define JET_REFINE_TOL 0.999
#endif

static long long median(std::vector<long long> v) {
    std::sort(v.begin(), v.end());
    size_t c = v.size();
    return c % 2 == 0 ? (v[c / 2] + v[c / 2 - 1]) / 2 : v[c / 2];
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "Insufficient number of args provided\n");
        std::fprintf(stderr, "Usage: %s <metis_graph_file> <config_file> <optional partition_output_filename> "
                             "<optional metrics_filename>\n", argv[0]);
        return -1;
    }
    jet_config_t cfg;
    if (jet_load_config(argv[2], &cfg) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        return -1;
    }
    cfg.refine_tolerance = JET_REFINE_TOL;
#ifdef JET_EXPORT
    cfg.dump_coarse = 1;
#endif
    jet_host_graph_t g;
    if (jet_load_metis(argv[1], &g) != JET_OK) {
        std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
        return -1;
    }
    std::printf("vertices: %lld; edges: %lld\n", (long long)g.n, (long long)(g.m / 2));
    const char* sd = std::getenv("JET_SEED");
    const unsigned long long seed = sd ? std::strtoull(sd, nullptr, 10) : 0ull;
    FILE* metrics = argc >= 5 ? std::fopen(argv[4], "w") : nullptr;
    std::vector<int32_t> part(g.n), best(g.n);
    std::vector<long long> cuts;
    long long best_cut = -1, sum = 0;
    const int attempts = cfg.num_iter;
    jet_config_t one = cfg;
    one.num_iter = 1;
    for (int i = 0; i < attempts; i++) {
        int64_t cut = 0;
        jet_stats_t st;
#ifdef JET_EXPORT
        // the reference driver's graph type (int32 offsets and weights) sets
        // the dump's widths
        std::vector<int32_t> rp32(g.row_offsets, g.row_offsets + g.n + 1);
        int rc = jet_partition_csr32(&one, g.n, rp32.data(), g.col, g.ew, nullptr, g.uniform_ew, JET_PTR_HOST,
                                     seed + (unsigned long long)i, part.data(), &cut, &st);
#else
        int rc = jet_partition_csr(&one, g.n, g.row_offsets, g.col, g.ew, nullptr, g.uniform_ew, JET_PTR_HOST,
                                   seed + (unsigned long long)i, part.data(), &cut, &st);
#endif
        if (rc != JET_OK) {
            std::fprintf(stderr, "FATAL ERROR: %s\n", jet_last_error());
            jet_free_graph(&g);
            return -1;
        }
        cuts.push_back(cut);
        sum += cut;
        if (best_cut < 0 || cut < best_cut) {
            best_cut = cut;
            best.swap(part);
        }
        if (metrics)
            std::fprintf(metrics,
                         "{\"attempt\": %d, \"edge_cut\": %lld, \"total_s\": %.6f, \"coarsen_s\": %.6f, "
                         "\"init_s\": %.6f, \"refine_s\": %.6f, \"levels\": %d, \"imbalance\": %.6f}\n",
                         i, (long long)cut, st.total_s, st.coarsen_s, st.init_partition_s, st.refine_s,
                         st.num_levels, st.finest_imb_ratio);
    }
    if (metrics) std::fclose(metrics);
    if (attempts > 0) {
        std::printf("graph %s, min edgecut found is %lld\n", argv[1], best_cut);
        std::printf("average edgecut: %lld\n", sum / attempts);
        std::printf("median edgecut: %lld\n", median(cuts));
        if (argc >= 4 && jet_write_partition(argv[3], g.n, best.data()) != JET_OK)
            std::fprintf(stderr, "could not write %s: %s\n", argv[3], jet_last_error());
    }
    jet_free_graph(&g);
    return 0;
}
