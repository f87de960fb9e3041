// Host copy of a coarse hierarchy read from a dump_coarse file
// (binary_dump.hpp:61-97): levels finest first, each with 64-bit row offsets,
// 32-bit columns and weights, and the map from the previous (finer) level.
#pragma once
#include <cstdint>
#include <vector>

struct jet_hierarchy {
    struct Level {
        int64_t n = 0, m = 0;
        std::vector<int64_t> rowptr;
        std::vector<int32_t> col, ew, vw;
        std::vector<int32_t> map;  // previous level -> this level (empty on level 0)
    };
    std::vector<Level> levels;
};
