// Process-wide tuning parameters (csrc/tuning.h) and their C-ABI accessors.
#include "tuning.h"

#include <cstdlib>
#include <cstring>
#include <mutex>

#include "jet_capi.h"

namespace jet {

Tuning& tuning() {
    static Tuning t;
    return t;
}

namespace {
struct Field {
    const char* name;
    int64_t Tuning::*member;
    int64_t lo, hi;
};
const Field kFields[] = {
    {"wg_max_n", &Tuning::wg_max_n, 0, 1 << 20},
};
std::mutex g_mu;
}  // namespace

}  // namespace jet

extern "C" int jet_tuning_set(const char* name, int64_t value) {
    if (!name) return JET_EINVAL;
    std::lock_guard<std::mutex> lk(jet::g_mu);
    for (const auto& f : jet::kFields) {
        if (std::strcmp(f.name, name) == 0) {
            if (value < f.lo || value > f.hi) return JET_EINVAL;
            jet::tuning().*(f.member) = value;
            return JET_OK;
        }
    }
    return JET_EINVAL;
}

extern "C" int jet_tuning_get(const char* name, int64_t* value) {
    if (!name || !value) return JET_EINVAL;
    std::lock_guard<std::mutex> lk(jet::g_mu);
    for (const auto& f : jet::kFields) {
        if (std::strcmp(f.name, name) == 0) {
            *value = jet::tuning().*(f.member);
            return JET_OK;
        }
    }
    return JET_EINVAL;
}
