// Per-kernel timing for measurement tools: HIP events on the library's stream
// around every launch of one kernel family, plus device-side item counters
// from which the family's algorithmic bytes are computed.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace jet {

enum ProfFamily : int { PF_NONE = 0, PF_LP_SELECT = 1, PF_CONN_BUILD = 2, PF_PROJECT = 3 };

struct Prof {
    int family = PF_NONE;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> pool;
    double host_bytes = 0;              // bytes known on the host (size-dependent terms)
    int64_t launches = 0;
    unsigned long long* dev = nullptr;  // device item counters (8)
};

Prof& prof();
inline bool prof_on(int fam) { return prof().family == fam; }
// Device counters when `fam` is being timed, else nullptr.
unsigned long long* prof_counters(int fam);
void prof_begin(int fam, hipStream_t s);
void prof_end(int fam, hipStream_t s, double host_bytes);
// Enables a family by name; returns false for an unknown name.
bool prof_enable(const std::string& name);
// Synchronizes, then returns launches, total milliseconds and algorithmic bytes.
bool prof_read(int64_t& launches, double& ms, double& bytes);

}  // namespace jet
