// Coarsening mappers on the GPU (heuristics.hpp of the reference).
#pragma once
#include "arena.h"
#include "jet_types.h"

namespace jet {

// Vertices whose degree exceeds this threshold are handled by one workgroup
// each instead of one lane group each.
constexpr int32_t HEAVY_DEG = 2048;

// Builds the list of heavy rows (degree > HEAVY_DEG); returns its length.
int32_t build_heavy_list(const DGraph& g, vid_t* list, Arena& a, hipStream_t s);

// HEC (heuristics.hpp:171-262): writes map[n], returns nc.
int32_t coarsen_hec(const DGraph& g, bool uniform, uint64_t seed, int32_t level, vid_t* map, Arena& a, hipStream_t s);

// Matching (heuristics.hpp:427-671); two_hop adds the leaf/twin/relative stages.
int32_t coarsen_match(const DGraph& g, bool uniform, uint64_t seed, int32_t level, bool two_hop, vid_t* map,
                      Arena& a, hipStream_t s);

// heuristics.hpp:89-169 on a given heavy-neighbour array; writes map, returns nc.
int32_t map_construct(vid_t n, const vid_t* hn, vid_t* map, Arena& a, hipStream_t s);

// HEC heavy-neighbour selection alone (heuristics.hpp:195-248).
void hec_heavy_neighbours(const DGraph& g, bool uniform, uint64_t seed, int32_t level, vid_t* hn, Arena& a,
                          hipStream_t s);

}  // namespace jet
