// Per-kernel-family timing and algorithmic-byte counting for measurement
// tools.
//
// Timing mode (one family at a time): every timed unit of the family -- one
// kernel launch, or for "update" one pass of update kernels -- is bracketed
// by two HIP events recorded on the stream the kernels are launched on; the
// family's time is the sum of the event intervals.
//
// Counting mode (any set of families, "all:count"): the units accumulate the
// item counts their algorithmic bytes follow from, per family: sizes the host
// knows (n, m) directly, data-dependent counts (cache misses and table slots
// read, candidate and move degrees) through device counters that are null
// outside counting runs. Counting adds work, so it runs in a separate pass of
// the same deterministic computation.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace jet {

enum ProfFamily : int { PF_LP_SELECT = 0, PF_AFTERBURNER, PF_UPDATE, PF_WEAK_SELECT, PF_FAMILIES };

// Family names as accepted by prof_enable, indexed by ProfFamily.
extern const char* const kProfNames[PF_FAMILIES];

constexpr int64_t PROF_MAX_UNITS = 32768;  // timed units per enable
constexpr int PROF_COUNTERS = 4;           // device counters per family

// Device counters of `fam` when it is being counted, else nullptr.
unsigned long long* prof_counters(int fam);
// Whether `fam` is being counted (host-side accounting is wanted).
bool prof_counting(int fam);
// Brackets one timed unit of `fam` on stream s (no-ops unless fam is timed).
void prof_begin(int fam, hipStream_t s);
void prof_end(int fam, hipStream_t s);
// Accounts one unit of `fam` and the bytes the host knows for it (counting).
void prof_account(int fam, double host_bytes);
// "lp_select" (etc.) times that family; "<family>:count" or "all:count"
// counts; "none" stops. Returns false for an unknown name.
bool prof_enable(const std::string& name);
// Synchronizes, then returns the timed family's units and total milliseconds
// (timing mode) or, for `fam`, its units and algorithmic bytes (counting).
bool prof_read(int fam, int64_t& units, double& ms, double& bytes);
// The family enabled for timing, or -1.
int prof_timed_family();

}  // namespace jet
