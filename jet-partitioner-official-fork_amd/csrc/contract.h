// Coarse-graph contraction (contract.hpp of the reference).
#pragma once
#include "arena.h"
#include "jet_types.h"

namespace jet {

// One level of the multilevel hierarchy (coarse_level_triple, contract.hpp:85-91).
struct Level {
    vid_t n = 0;
    eoff_t m = 0;
    eoff_t* rowptr = nullptr;
    vid_t* col = nullptr;
    wgt_t* ew = nullptr;
    wgt_t* vw = nullptr;
    vid_t* map = nullptr;  // previous (finer) level -> this level
    bool uniform = false;
    vid_t* heavy = nullptr;  // rows with degree > HEAVY_DEG
    int32_t nheavy = 0;
    DGraph view() const {
        DGraph g;
        g.n = n;
        g.m = m;
        g.rowptr = rowptr;
        g.col = col;
        g.ew = ew;
        g.vw = vw;
        return g;
    }
};

// Builds the coarse level for `map` (n -> nc). The coarse arrays come from
// `persist`, temporaries from `a`.
void contract(const DGraph& g, const vid_t* map, vid_t nc, Arena& persist, Arena& a, hipStream_t s, Level& out);

}  // namespace jet
