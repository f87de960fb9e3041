// Per-kernel timing for measurement tools. When a kernel family is enabled,
// each launch of it receives a span slot: its first workgroup to start and its
// last to finish stamp the device's constant-rate wall clock there (atomic min
// and max), so the span is the kernel's own execution time, free of host
// submission gaps and of the stream's marker packets. Device item counters
// give the family's algorithmic bytes.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace jet {

enum ProfFamily : int { PF_NONE = 0, PF_LP_SELECT = 1, PF_CONN_BUILD = 2, PF_PROJECT = 3 };

constexpr int64_t PROF_MAX_SPANS = 65536;

struct Prof {
    int family = PF_NONE;
    double host_bytes = 0;              // bytes known on the host (size-dependent terms)
    int64_t launches = 0;               // launches accounted
    int64_t spans = 0;                  // span slots handed out (<= PROF_MAX_SPANS)
    unsigned long long* dev = nullptr;  // device item counters (8)
    unsigned long long* span = nullptr; // device [start, end] per launch
};

Prof& prof();
inline bool prof_on(int fam) { return prof().family == fam; }
// Device counters when `fam` is being timed, else nullptr.
unsigned long long* prof_counters(int fam);
// Span slot (two wall-clock words) for the next launch of `fam`, or nullptr.
unsigned long long* prof_span(int fam);
// Accounts one launch of `fam`.
void prof_account(int fam, double host_bytes);
// Enables a family by name; returns false for an unknown name.
bool prof_enable(const std::string& name);
// Synchronizes, then returns launches, total milliseconds and algorithmic bytes.
bool prof_read(int64_t& launches, double& ms, double& bytes);

// Device side: call from every thread at kernel entry and after the kernel's
// last store (the latter behind a workgroup barrier).
__device__ inline void prof_span_begin(unsigned long long* sp) {
    if (sp && threadIdx.x == 0) atomicMin(&sp[0], (unsigned long long)wall_clock64());
}
__device__ inline void prof_span_end(unsigned long long* sp) {
    if (sp && threadIdx.x == 0) {
        __threadfence();
        atomicMax(&sp[1], (unsigned long long)wall_clock64());
    }
}

}  // namespace jet
