// Host initial partitioner (csrc/initpart.cpp): the C entry point for the
// coarsest graph, and the coarsening hierarchy it uses for large graphs.
#pragma once
#include <cstdint>
#include <vector>

extern "C" int jet_host_initpart(int32_t n, const int64_t* xadj, const int32_t* adj, const int32_t* adjw,
                                 const int32_t* vw, int32_t k, double imb, uint64_t seed, int32_t* part);

namespace jet_init {

// One coarser level of the initial partitioner's hierarchy.
struct HostLevel {
    int32_t n = 0;
    std::vector<int64_t> xadj;
    std::vector<int32_t> adj, adjw, vw;
    std::vector<int32_t> map;  // vertex of the next finer level -> vertex of this level
};

// Coarsest graphs with at least this many adjacency entries are partitioned
// multilevel k-way: coarsened further on the host (init_hierarchy), the
// smallest level partitioned by jet_host_initpart, and the partition refined
// back up with Jet refinement level by level (the structure of METIS's k-way
// scheme, with Jet as its refiner).
constexpr int64_t INIT_ML_EDGES = int64_t(1) << 20;

inline bool wants_hierarchy(int32_t n, int64_t m, int32_t k) {
    return m >= INIT_ML_EDGES && (int64_t)n > 8 * (int64_t)k;
}

// Caps the host threads that the initial partitioner uses for calls made on
// the calling thread (0: no cap). Results do not depend on the thread count.
void set_thread_budget(int threads);

// The levels coarser than the graph (first = one level coarser), down to
// about max(1000, 8 k) vertices or until a level shrinks by less than 10%.
std::vector<HostLevel> init_hierarchy(int32_t n, const int64_t* xadj, const int32_t* adj, const int32_t* adjw,
                                      const int32_t* vw, int32_t k, uint64_t seed);

}  // namespace jet_init
