// Per-kernel-family timing and algorithmic-byte counting for measurement
// tools.
//
// Timing mode (one family at a time): every kernel launch of the family is
// made with hipExtLaunchKernelGGL and a pair of HIP events that the launch
// itself records at the kernel's start and end, after a marker event on the
// same stream; an interval runs from the later of the marker and the start
// stamp to the end stamp, so it is the kernel's execution alone (what a
// profiler's dispatch record measures), neither the launch gap before it
// nor, for a launch queued behind a running kernel (whose start stamp can be
// taken early), the wait for that kernel. The family's time is the sum over
// its launches.
//
// Counting mode (any set of families, "all:count"): the launches are counted
// and the item counts their algorithmic bytes follow from are accumulated,
// per family: sizes the host knows (n, m) directly, data-dependent counts
// (cache misses and table slots read, candidate and move degrees) through
// device counters that are null outside counting runs. Counting adds work, so
// it runs in a separate pass of the same deterministic computation.
#pragma once
#include <hip/hip_runtime.h>

#include <hip/hip_ext.h>

#include <string>
#include <utility>

namespace jet {

enum ProfFamily : int { PF_LP_SELECT = 0, PF_AFTERBURNER, PF_UPDATE, PF_WEAK_SELECT, PF_FAMILIES };

// Family names as accepted by prof_enable, indexed by ProfFamily.
extern const char* const kProfNames[PF_FAMILIES];

constexpr int64_t PROF_MAX_UNITS = 65536;  // timed launches per enable
constexpr int PROF_COUNTERS = 4;           // device counters per family

// Device counters of `fam` when it is being counted, else nullptr.
unsigned long long* prof_counters(int fam);
// Whether `fam` is being counted (host-side accounting is wanted).
bool prof_counting(int fam);
// Start and end events for one launch of `fam` on stream s: a fresh pair
// when fam is timed (after a marker event recorded on s), null events
// otherwise; in counting mode the launch is counted.
struct ProfEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
ProfEvents prof_launch(int fam, hipStream_t s);
// Adds bytes the host knows to `fam` (counting mode; no-op otherwise).
void prof_bytes(int fam, double host_bytes);
// "lp_select" (etc.) times that family; "<family>:count" or "all:count"
// counts; "none" stops. Returns false for an unknown name.
bool prof_enable(const std::string& name);
// Synchronizes, then returns the timed family's launches and total kernel
// milliseconds (timing mode) or, for `fam`, its launches and algorithmic
// bytes (counting).
bool prof_read(int fam, int64_t& units, double& ms, double& bytes);
// The family enabled for timing, or -1.
int prof_timed_family();

// A launch of a kernel of family `fam` on stream s: with its own start and
// end events when the family is timed, counted when it is counted.
template <typename... P, typename... A>
inline void family_launch(int fam, void (*kernel)(P...), dim3 grid, dim3 block, uint32_t shm, hipStream_t s,
                          A&&... args) {
    const ProfEvents e = prof_launch(fam, s);
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, e.start, e.stop, 0, std::forward<A>(args)...);
}

}  // namespace jet
