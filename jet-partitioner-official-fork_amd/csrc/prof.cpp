// Kernel-family timing (HIP events) and algorithmic-byte counting (prof.h).
#include "prof.h"

#include <cstring>
#include <vector>

#include "jet_types.h"

namespace jet {

const char* const kProfNames[PF_FAMILIES] = {"lp_select", "afterburner", "update", "weak_select"};

namespace {

struct Prof {
    int timed = -1;                       // family timed with events
    bool counting[PF_FAMILIES] = {};      // families counted
    double host_bytes[PF_FAMILIES] = {};  // bytes known on the host
    int64_t units[PF_FAMILIES] = {};      // launches counted (counting)
    unsigned long long* dev = nullptr;    // device counters, PROF_COUNTERS per family
    std::vector<hipEvent_t> ev;           // 3 per timed launch: marker, start, end
    int64_t nev = 0;                      // timed launches made
    int dev_id = -1;                      // device the events and counters belong to
};

Prof& P() {
    static Prof* p = new Prof();
    return *p;
}

int family_of(const std::string& name) {
    for (int f = 0; f < PF_FAMILIES; f++)
        if (name == kProfNames[f]) return f;
    return -1;
}

// Algorithmic bytes of one family from its counters, at the declared widths
// (ordinal 4, offset 8, weight / connectivity / gain 4, part 4).
double counted_bytes(int fam, const unsigned long long* c) {
    switch (fam) {
        case PF_LP_SELECT:
            // 8 per vertex (cached destination, lock) + 28 per re-evaluated
            // vertex (offset 8, size 4, part 4, gain 4, destination 4, cache 4)
            // + 8 per table slot read; c = {evaluated, slots}
            return 28.0 * (double)c[0] + 8.0 * (double)c[1];
        case PF_AFTERBURNER:
            // 16 per candidate edge (column 4, weight 4, neighbour's
            // destination 4 and pre-gain 4); c = {sum of candidate degrees}
            return 16.0 * (double)c[0];
        case PF_UPDATE:
            // per pass, 24 per move edge (column 4, weight 4, row offset 8,
            // table slot 8); c = {sum of move degrees over both passes}
            return 24.0 * (double)c[0];
        case PF_WEAK_SELECT:
            // 12 per vertex (part, destination, bucket id) + 16 per vertex of
            // an oversized part (weight 4, offset 8, size 4) + 8 per table
            // slot read; c = {evaluated, slots}
            return 16.0 * (double)c[0] + 8.0 * (double)c[1];
        default:
            return 0;
    }
}

}  // namespace

unsigned long long* prof_counters(int fam) {
    Prof& p = P();
    return fam >= 0 && fam < PF_FAMILIES && p.counting[fam] && p.dev ? p.dev + PROF_COUNTERS * fam : nullptr;
}

bool prof_counting(int fam) { return fam >= 0 && fam < PF_FAMILIES && P().counting[fam]; }

ProfEvents prof_launch(int fam, hipStream_t s) {
    Prof& p = P();
    ProfEvents e;
    if (fam < 0 || fam >= PF_FAMILIES) return e;
    if (p.counting[fam]) p.units[fam]++;
    if (p.timed != fam || p.nev >= PROF_MAX_UNITS) return e;
    // a marker first: it completes when the work queued before the launch
    // has finished, which bounds the kernel's start from below when the
    // launch's own start stamp is taken while that work still runs
    JET_HIP(hipEventRecord(p.ev[3 * p.nev], s));
    e.start = p.ev[3 * p.nev + 1];
    e.stop = p.ev[3 * p.nev + 2];
    p.nev++;
    return e;
}

void prof_bytes(int fam, double host_bytes) {
    Prof& p = P();
    if (fam < 0 || fam >= PF_FAMILIES || !p.counting[fam]) return;
    p.host_bytes[fam] += host_bytes;
}

int prof_timed_family() { return P().timed; }

bool prof_enable(const std::string& spec) {
    Prof& p = P();
    std::string name = spec;
    bool counting = false;
    const std::string suffix = ":count";
    if (name.size() > suffix.size() && name.compare(name.size() - suffix.size(), suffix.size(), suffix) == 0) {
        counting = true;
        name.resize(name.size() - suffix.size());
    }
    int fam = -1;
    if (name != "none" && !(counting && name == "all")) {
        fam = family_of(name);
        if (fam < 0) return false;
    }
    JET_HIP(hipDeviceSynchronize());
    int dev = 0;
    JET_HIP(hipGetDevice(&dev));
    if (p.dev_id != dev) {  // counters and events of the current device
        p.dev = nullptr;
        p.ev.clear();
        p.dev_id = dev;
    }
    p.timed = -1;
    std::memset(p.counting, 0, sizeof(p.counting));
    std::memset(p.host_bytes, 0, sizeof(p.host_bytes));
    std::memset(p.units, 0, sizeof(p.units));
    p.nev = 0;
    if (name == "none") return true;
    if (counting) {
        if (!p.dev) JET_HIP(hipMalloc(&p.dev, sizeof(unsigned long long) * PROF_COUNTERS * PF_FAMILIES));
        JET_HIP(hipMemset(p.dev, 0, sizeof(unsigned long long) * PROF_COUNTERS * PF_FAMILIES));
        JET_HIP(hipDeviceSynchronize());
        for (int f = 0; f < PF_FAMILIES; f++) p.counting[f] = fam < 0 || f == fam;
        return true;
    }
    if (p.ev.empty()) {
        p.ev.resize(3 * PROF_MAX_UNITS);
        for (auto& e : p.ev) JET_HIP(hipEventCreate(&e));
    }
    p.timed = fam;
    return true;
}

bool prof_read(int fam, int64_t& units, double& ms, double& bytes) {
    Prof& p = P();
    JET_HIP(hipDeviceSynchronize());
    units = 0;
    ms = 0;
    bytes = 0;
    if (fam < 0 || fam >= PF_FAMILIES) return false;
    if (p.counting[fam]) {
        unsigned long long c[PROF_COUNTERS];
        JET_HIP(hipMemcpy(c, p.dev + PROF_COUNTERS * fam, sizeof(c), hipMemcpyDeviceToHost));
        units = p.units[fam];
        bytes = p.host_bytes[fam] + counted_bytes(fam, c);
        return true;
    }
    if (p.timed != fam) return false;
    for (int64_t i = 0; i < p.nev; i++) {
        // from the later of the marker and the launch's start stamp
        float t_start = 0, t_marker = 0;
        JET_HIP(hipEventElapsedTime(&t_start, p.ev[3 * i + 1], p.ev[3 * i + 2]));
        JET_HIP(hipEventElapsedTime(&t_marker, p.ev[3 * i], p.ev[3 * i + 2]));
        ms += t_start < t_marker ? t_start : t_marker;
    }
    units = p.nev;
    return true;
}

}  // namespace jet
