// Per-kernel timing for measurement tools: HIP events bound to every launch of
// one kernel family on the library's stream, plus device-side item counters
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
// Event pair for the next launch of `fam` ({nullptr, nullptr} when the family
// is not being timed). The launch passes them to hipExtLaunchKernelGGL, which
// stamps them at the kernel's own start and end, so host submission gaps on
// an idle stream are not counted.
std::pair<hipEvent_t, hipEvent_t> prof_events(int fam);
// Accounts one launch of `fam` whose events came from prof_events.
void prof_account(int fam, double host_bytes);
// Enables a family by name; returns false for an unknown name.
bool prof_enable(const std::string& name);
// Synchronizes, then returns launches, total milliseconds and algorithmic bytes.
bool prof_read(int64_t& launches, double& ms, double& bytes);

}  // namespace jet
