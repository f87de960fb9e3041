// Per-kernel timing for measurement tools. When a kernel family is enabled for
// timing, each launch of it receives a span record: its workgroups stamp the
// device's constant-rate wall clock at entry (atomic min) and after their last
// store (atomic max), spread over PROF_SPAN_SLOTS words each so the atomics
// stay cheap; the launch's span is the earliest start to the latest end, the
// kernel's own execution time without host submission gaps. In counting mode
// the family's launches instead accumulate the item counts from which its
// algorithmic bytes follow; counting adds work to the kernel, so the two modes
// are used in separate runs of the same deterministic computation.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace jet {

enum ProfFamily : int { PF_NONE = 0, PF_LP_SELECT = 1 };

constexpr int PROF_SPAN_SLOTS = 64;
constexpr int64_t PROF_MAX_SPANS = 16384;

struct Prof {
    int family = PF_NONE;
    bool counting = false;
    double host_bytes = 0;               // bytes known on the host (size-dependent terms)
    int64_t launches = 0;                // launches accounted
    int64_t spans = 0;                   // span records handed out (<= PROF_MAX_SPANS)
    unsigned long long* dev = nullptr;   // device item counters (8)
    unsigned long long* span = nullptr;  // per launch: starts[SLOTS], ends[SLOTS]
};

Prof& prof();
inline bool prof_on(int fam) { return prof().family == fam; }
// Device counters when `fam` is being counted, else nullptr.
unsigned long long* prof_counters(int fam);
// Span record for the next launch of `fam` when it is being timed, else nullptr.
unsigned long long* prof_span(int fam);
// Accounts one launch of `fam`.
void prof_account(int fam, double host_bytes);
// Enables a family by name ("lp_select" times it, "lp_select:count" counts
// its items, "none" stops); returns false for an unknown name.
bool prof_enable(const std::string& name);
// Synchronizes, then returns launches, total milliseconds (timing mode) and
// algorithmic bytes (counting mode; the host-known part in timing mode).
bool prof_read(int64_t& launches, double& ms, double& bytes);

// Device side: call from every thread at kernel entry, and after the kernel's
// last store behind a workgroup barrier.
__device__ inline void prof_span_begin(unsigned long long* sp) {
    if (sp && threadIdx.x == 0) atomicMin(&sp[blockIdx.x % PROF_SPAN_SLOTS], (unsigned long long)wall_clock64());
}
__device__ inline void prof_span_end(unsigned long long* sp) {
    if (sp && threadIdx.x == 0)
        atomicMax(&sp[PROF_SPAN_SLOTS + blockIdx.x % PROF_SPAN_SLOTS], (unsigned long long)wall_clock64());
}

}  // namespace jet
