// Launch-shape and path parameters that do not change results (every setting
// gives the same partition), read at context creation and settable through
// jet_tuning_set for the autotuner and for tests.
#pragma once
#include <cstdint>

namespace jet {

struct Tuning {
    // levels of up to this many vertices are refined by the one-workgroup
    // kernel (0: never)
    int64_t wg_max_n = 16384;
};

Tuning& tuning();

}  // namespace jet
